// qg_hpsi.hip — the H·psi hot path on MI355X (gfx950): the fused 8th-order Laplacian +
// local potential + Kleinman-Bylander projector apply over a block of bands, the projector
// dot-product kernel, and their C-ABI entry points.
//
// Data layout in HBM: band-major blocks, x fastest (psi[b*ld + ix + Nx*(iy + Ny*iz)]).
//
// Fused kernel design (one pass over psi and Hpsi):
//   * a workgroup owns TY full x-rows of the grid, a slab of TZ z-planes and NB bands;
//   * it streams through its slab in z, keeping each point's 9 z-neighbours in registers
//     (a register queue), and stages the current plane of every band with a 4-row y-halo
//     in LDS (double-buffered, one barrier per plane); x-neighbours wrap inside the LDS
//     row because the tile spans the whole periodic x extent, so there is no x-halo;
//   * V and the projector terms are applied in the same pass: per (y tile, z plane) the
//     host precomputes the list of projector images that reach it, each image carries
//     exact per-row [x0,x1) sphere intervals, and beta = p(r) Y_lm is evaluated in-kernel
//     (no per-point projector tables in HBM);
//   * workgroup ids are remapped so that consecutive tiles run on the same XCD and share
//     that XCD's L2 for the halo rows.
// Projector coefficients come from a separate atom-centric kernel (one workgroup per
// atom and band chunk, rows of the sphere mapped to half-waves, wavefront reductions).
#include <algorithm>
#include <cmath>
#include "qg_internal.hpp"

namespace qg {

constexpr int kBlock = 256;
constexpr double kY00 = 0.28209479177387814;     // 1/sqrt(4 pi)
constexpr double kY1 = 0.48860251190291992;      // sqrt(3/(4 pi))

template <class T>
struct StencilCoef {
    T c0;
    T cx[4], cy[4], cz[4];
};

enum { MODE_LAP = 0, MODE_HAM = 1, MODE_NL = 2 };

struct FusedArgs {
    int nx, ny, nz;
    int ty, tz;
    int ntile_y, nslab;
    int nbands;
    int64_t ld_in, ld_out;
    double h0, h1, h2;
    int accumulate;
    // nonlocal (MODE_HAM with projectors, MODE_NL)
    int has_nl;
    const int* list_off;
    const int* list_ids;
    const DevImage* images;
    const DevProj* projs;
    const double* tables;
    const int2* intervals;
    const double* coef;
    int ncoef;
};

// Map the hardware workgroup id (dispatched round-robin over the 8 XCDs) to a logical id
// such that each XCD receives one contiguous range of logical ids.
__device__ __forceinline__ int xcd_remap(int i, int n) {
    const int xcd = i & 7, slot = i >> 3;
    const int per = n >> 3, rem = n & 7;
    const int start = (xcd < rem) ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    return start + slot;
}

// RadialGrid2::GetValueByXi (RadialGrid2.h:81-136) on the device (FP64).
__device__ double radial_by_xi(double xi, const double* v, int n, double xi_min, double xi_delta) {
    const double rel = xi - xi_min;
    const double fl = floor(rel / xi_delta);
    const int i = (int)fl;
    if (i >= n - 1) return 0.0;
    if (i == 0 || i == n - 2) {
        const double t = rel / xi_delta - fl;
        return v[i] * (1.0 - t) + v[i + 1] * t;
    }
    const double t = rel / xi_delta - (double)i;
    if (1 < i && i < n - 3) {
        const double fm2 = v[i - 2], fm1 = v[i - 1], f0 = v[i], f1 = v[i + 1], f2 = v[i + 2], f3 = v[i + 3];
        const double g0 = 2.0 / 3.0 * (f1 - fm1) - 1.0 / 12.0 * (f2 - fm2);
        const double g1 = 2.0 / 3.0 * (f2 - f0) - 1.0 / 12.0 * (f3 - fm1);
        const double h0 = 4.0 / 3.0 * (f1 + fm1) - 1.0 / 12.0 * (f2 + fm2) - 5.0 / 2.0 * f0;
        const double h1 = 4.0 / 3.0 * (f2 + f0) - 1.0 / 12.0 * (f3 + fm1) - 5.0 / 2.0 * f1;
        const double e0 = f1 - (f0 + g0 + h0 / 2.0);
        const double e1 = g1 - (g0 + h0);
        const double e2 = h1 - h0;
        const double a3 = 10.0 * e0 - 4.0 * e1 + 0.5 * e2;
        const double a4 = -15.0 * e0 + 7.0 * e1 - e2;
        const double a5 = f1 - f0 - g0 - h0 / 2.0 - a3 - a4;
        return f0 + t * (g0 + t * (h0 / 2.0 + t * (a3 + t * (a4 + t * a5))));
    }
    const double v0 = v[i], v1 = v[i + 1];
    const double s1 = (v[i + 2] - v0) / 2.0, s0 = (v1 - v[i - 1]) / 2.0;
    const double a = s1 + s0 - 2.0 * (v1 - v0);
    const double b = -s1 - 2.0 * s0 + 3.0 * (v1 - v0);
    return a * t * t * t + b * t * t + s0 * t + v0;
}

// Radial projector value p(r) at squared distance rr (inside the cutoff by construction).
template <class T>
__device__ __forceinline__ T radial_value(const DevProj& p, const double* tables, double rr) {
    if (p.kind == QG_RADIAL_GAUSSIAN) {
        if constexpr (sizeof(T) == 8) {
            return (T)(p.amp * exp(-rr * p.inv_2s2));
        } else {
            return (T)p.amp * expf(-(float)rr * (float)p.inv_2s2);
        }
    }
    const double* t = tables + p.table_off;
    if (p.r_min2 > rr) return (T)t[0];
    return (T)radial_by_xi(0.5 * log(rr), t, p.n_table, p.xi_min, p.xi_delta);
}

// Offsets of grid point (ix,iy,iz) from an image's nucleus (reference arithmetic order).
__device__ __forceinline__ void image_offset(const DevImage& im, double h0, double h1, double h2,
                                             int ix, int iy, int iz, double& x, double& y, double& z,
                                             double& rr) {
    z = __dsub_rn(__dmul_rn(h2, (double)(iz - im.shift[2])), im.R[2]);
    y = __dsub_rn(__dmul_rn(h1, (double)(iy - im.shift[1])), im.R[1]);
    x = __dsub_rn(__dmul_rn(h0, (double)(ix - im.shift[0])), im.R[0]);
    const double zz = __dmul_rn(z, z);
    const double yy_zz = __dadd_rn(__dmul_rn(y, y), zz);
    rr = __dadd_rn(__dmul_rn(x, x), yy_zz);
}

// ---------------------------------------------------------------------------------------
// Fused kernel.  T = float/double, NB bands per workgroup, PPT points per thread.
template <class T, int NB, int PPT, int MODE>
__global__ __launch_bounds__(kBlock) void fused_hpsi_kernel(FusedArgs a, StencilCoef<T> sc,
                                                           const T* __restrict__ V,
                                                           const T* __restrict__ psi,
                                                           T* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int lds_band = (a.ty + 8) * nx;

    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const int ytile = wg % a.ntile_y;
    const int rest = wg / a.ntile_y;
    const int slab = rest % a.nslab;
    const int chunk = rest / a.nslab;
    const int y0 = ytile * a.ty;
    const int rows = min(a.ty, ny - y0);
    const int z0 = slab * a.tz;
    const int z1 = min(z0 + a.tz, nz);
    const int b0 = chunk * NB;
    const int tid = threadIdx.x;
    const int npts = rows * nx;
    const int64_t plane = (int64_t)nx * ny;
    const int64_t tile_off = (int64_t)y0 * nx;

    int pix[PPT], piy[PPT];
    bool pv[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int p = tid + kBlock * k;
        pv[k] = p < npts;
        piy[k] = pv[k] ? p / nx : 0;
        pix[k] = pv[k] ? p - piy[k] * nx : 0;
    }
    const T* src[NB];
    T* dst[NB];
    bool bv[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        bv[b] = (b0 + b) < a.nbands;
        const int bb = bv[b] ? b0 + b : a.nbands - 1;
        src[b] = psi + (int64_t)bb * a.ld_in;
        dst[b] = out + (int64_t)bb * a.ld_out;
    }
    // halo elements of this thread: 8 rows x nx (4 above, 4 below the tile), wrapped in y
    constexpr int HPPT = 4;   // supports nx <= 128
    int h_lds[HPPT], h_glb[HPPT];
    bool hv[HPPT];
#pragma unroll
    for (int j = 0; j < HPPT; ++j) {
        const int e = tid + kBlock * j;
        hv[j] = (MODE != MODE_NL) && e < 8 * nx;
        const int r8 = hv[j] ? e / nx : 0;
        const int col = hv[j] ? e - r8 * nx : 0;
        int gy = (r8 < 4) ? y0 - 4 + r8 : y0 + rows + (r8 - 4);
        gy = ((gy % ny) + ny) % ny;
        const int ly = (r8 < 4) ? r8 : rows + 4 + (r8 - 4);
        h_lds[j] = ly * nx + col;
        h_glb[j] = gy * nx + col;
    }
    auto zwrap = [&](int z) { int zz = z % nz; return zz < 0 ? zz + nz : zz; };

    T q[NB][PPT][9];
    if constexpr (MODE != MODE_NL) {
#pragma unroll
        for (int d = 0; d < 8; ++d) {
            const int64_t off = (int64_t)zwrap(z0 - 4 + d) * plane + tile_off;
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int k = 0; k < PPT; ++k) q[b][k][d] = pv[k] ? src[b][off + tid + kBlock * k] : T(0);
        }
    }

    for (int z = z0; z < z1; ++z) {
        const int64_t zoff = (int64_t)z * plane;
        T res[NB][PPT];
        if constexpr (MODE != MODE_NL) {
            const int64_t off = (int64_t)zwrap(z + 4) * plane + tile_off;
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int k = 0; k < PPT; ++k) q[b][k][8] = pv[k] ? src[b][off + tid + kBlock * k] : T(0);
            T* L = lds + ((z - z0) & 1) * NB * lds_band;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
#pragma unroll
                for (int k = 0; k < PPT; ++k)
                    if (pv[k]) L[b * lds_band + 4 * nx + tid + kBlock * k] = q[b][k][4];
#pragma unroll
                for (int j = 0; j < HPPT; ++j)
                    if (hv[j]) L[b * lds_band + h_lds[j]] = src[b][zoff + h_glb[j]];
            }
        }
        // local potential / accumulate base
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            T v = T(0);
            if constexpr (MODE == MODE_HAM) v = pv[k] ? V[zoff + tile_off + tid + kBlock * k] : T(0);
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if constexpr (MODE == MODE_HAM) {
                    res[b][k] = v * q[b][k][4];
                } else {
                    res[b][k] = (a.accumulate && pv[k] && bv[b]) ? dst[b][zoff + tile_off + tid + kBlock * k] : T(0);
                }
            }
        }
        // nonlocal projector terms
        if (MODE != MODE_LAP && a.has_nl) {
            const int slot = ytile * nz + z;
            const int lo = a.list_off[slot], hi = a.list_off[slot + 1];
            for (int e = lo; e < hi; ++e) {
                const DevImage im = a.images[a.list_ids[e]];
                const int wy = im.e[1] - im.b[1];
                const bool zin = (z >= im.b[2]) && (z < im.e[2]);
                if (!zin) continue;
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    const int iy = y0 + piy[k];
                    const int ix = pix[k];
                    bool in = pv[k] && iy >= im.b[1] && iy < im.e[1];
                    if (in) {
                        const int2 iv = a.intervals[im.row_base + (iy - im.b[1]) + wy * (z - im.b[2])];
                        in = ix >= iv.x && ix < iv.y;
                    }
                    if (!in) continue;
                    double x, y, zc, rr;
                    image_offset(im, a.h0, a.h1, a.h2, ix, iy, z, x, y, zc, rr);
                    for (int pj = im.proj_begin; pj < im.proj_end; ++pj) {
                        const DevProj pr = a.projs[pj];
                        const T rad = radial_value<T>(pr, a.tables, rr);
                        const double* cb = a.coef + im.coef_base + pr.coef_rel;
                        if (pr.l == 0) {
                            const T beta = rad * (T)kY00;
#pragma unroll
                            for (int b = 0; b < NB; ++b)
                                res[b][k] += (T)(pr.energy * cb[(int64_t)(b0 + (bv[b] ? b : 0)) * a.ncoef]) * beta;
                        } else {
                            T ym1 = T(0), y0v = T(0), yp1 = T(0);
                            if (!(1.0e-14 > rr)) {
                                const double ir = 1.0 / sqrt(rr);
                                ym1 = (T)(kY1 * (y * ir));
                                y0v = (T)(kY1 * (zc * ir));
                                yp1 = (T)(kY1 * (x * ir));
                            }
#pragma unroll
                            for (int b = 0; b < NB; ++b) {
                                const double* c = cb + (int64_t)(b0 + (bv[b] ? b : 0)) * a.ncoef;
                                res[b][k] += (T)(pr.energy * c[0]) * (rad * ym1) +
                                             (T)(pr.energy * c[1]) * (rad * y0v) +
                                             (T)(pr.energy * c[2]) * (rad * yp1);
                            }
                        }
                    }
                }
            }
        }
        if constexpr (MODE != MODE_NL) {
            __syncthreads();
            const T* L = lds + ((z - z0) & 1) * NB * lds_band;
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                if (!pv[k]) continue;
                const int ix = pix[k];
                const int c = (piy[k] + 4) * nx + ix;
                int xm[4], xp[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    int t = ix - (m + 1);
                    xm[m] = (t < 0 ? t + nx : t) - ix;
                    t = ix + (m + 1);
                    xp[m] = (t >= nx ? t - nx : t) - ix;
                }
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const T* Lb = L + b * lds_band + c;
                    T d = sc.c0 * q[b][k][4];
                    d += sc.cx[3] * (Lb[xp[3]] + Lb[xm[3]]) + sc.cx[2] * (Lb[xp[2]] + Lb[xm[2]]) +
                         sc.cx[1] * (Lb[xp[1]] + Lb[xm[1]]) + sc.cx[0] * (Lb[xp[0]] + Lb[xm[0]]);
                    d += sc.cy[3] * (Lb[4 * nx] + Lb[-4 * nx]) + sc.cy[2] * (Lb[3 * nx] + Lb[-3 * nx]) +
                         sc.cy[1] * (Lb[2 * nx] + Lb[-2 * nx]) + sc.cy[0] * (Lb[nx] + Lb[-nx]);
                    d += sc.cz[3] * (q[b][k][8] + q[b][k][0]) + sc.cz[2] * (q[b][k][7] + q[b][k][1]) +
                         sc.cz[1] * (q[b][k][6] + q[b][k][2]) + sc.cz[0] * (q[b][k][5] + q[b][k][3]);
                    res[b][k] += d;
                }
            }
        }
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int k = 0; k < PPT; ++k)
                if (pv[k] && bv[b]) dst[b][zoff + tile_off + tid + kBlock * k] = res[b][k];
        if constexpr (MODE != MODE_NL) {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int k = 0; k < PPT; ++k)
#pragma unroll
                    for (int d = 0; d < 8; ++d) q[b][k][d] = q[b][k][d + 1];
        }
    }
}

// ---------------------------------------------------------------------------------------
// Projector coefficients: coef[b*ncoef + j] = dV * sum_i beta_j(i) psi_b(i).
// One workgroup per (atom, chunk of NBA bands); the atom's image rows are spread over the
// 8 half-waves of the block, lanes walk x inside the exact sphere interval.
struct ProjArgs {
    int nx, ny, nz;
    int natoms;
    int nbands;
    int64_t ld;
    double h0, h1, h2, dv;
    const int* atom_img_off;
    const int* atom_coef_off;
    const DevImage* images;
    const DevProj* projs;
    const double* tables;
    const int2* intervals;
    double* coef;
    int ncoef;
};

template <int MAXJ, class T>
__device__ __forceinline__ void eval_beta(const DevImage& im, const DevProj* projs, const double* tables,
                                          double x, double y, double z, double rr, T beta[MAXJ]) {
#pragma unroll
    for (int j = 0; j < MAXJ; ++j) beta[j] = T(0);
    for (int pj = im.proj_begin; pj < im.proj_end; ++pj) {
        const DevProj pr = projs[pj];
        const T rad = radial_value<T>(pr, tables, rr);
        if (pr.l == 0) {
#pragma unroll
            for (int j = 0; j < MAXJ; ++j)
                if (j == pr.coef_rel) beta[j] = rad * (T)kY00;
        } else {
            T ym1 = T(0), y0v = T(0), yp1 = T(0);
            if (!(1.0e-14 > rr)) {
                const double ir = 1.0 / sqrt(rr);
                ym1 = (T)(kY1 * (y * ir));
                y0v = (T)(kY1 * (z * ir));
                yp1 = (T)(kY1 * (x * ir));
            }
#pragma unroll
            for (int j = 0; j < MAXJ; ++j) {
                if (j == pr.coef_rel) beta[j] = rad * ym1;
                if (j == pr.coef_rel + 1) beta[j] = rad * y0v;
                if (j == pr.coef_rel + 2) beta[j] = rad * yp1;
            }
        }
    }
}

template <class T, int NBA, int MAXJ>
__global__ __launch_bounds__(kBlock) void project_kernel(ProjArgs a, const T* __restrict__ psi) {
    using Acc = T;
    __shared__ double red[kBlock / 64][MAXJ * NBA];
    const int atom = blockIdx.x % a.natoms;
    const int chunk = blockIdx.x / a.natoms;
    const int b0 = chunk * NBA;
    const int tid = threadIdx.x;
    const int half = tid >> 5, lane = tid & 31;
    const int64_t plane = (int64_t)a.nx * a.ny;
    const T* src[NBA];
#pragma unroll
    for (int b = 0; b < NBA; ++b) src[b] = psi + (int64_t)min(b0 + b, a.nbands - 1) * a.ld;
    Acc acc[MAXJ][NBA];
#pragma unroll
    for (int j = 0; j < MAXJ; ++j)
#pragma unroll
        for (int b = 0; b < NBA; ++b) acc[j][b] = Acc(0);

    for (int ii = a.atom_img_off[atom]; ii < a.atom_img_off[atom + 1]; ++ii) {
        const DevImage im = a.images[ii];
        const int wy = im.e[1] - im.b[1];
        const int nrows = wy * (im.e[2] - im.b[2]);
        for (int row = half; row < nrows; row += kBlock / 32) {
            const int iz = im.b[2] + row / wy;
            const int iy = im.b[1] + row % wy;
            const int2 iv = a.intervals[im.row_base + row];
            for (int ix = iv.x + lane; ix < iv.y; ix += 32) {
                double x, y, z, rr;
                image_offset(im, a.h0, a.h1, a.h2, ix, iy, iz, x, y, z, rr);
                T beta[MAXJ];
                eval_beta<MAXJ, T>(im, a.projs, a.tables, x, y, z, rr, beta);
                const int64_t gi = (int64_t)iz * plane + (int64_t)iy * a.nx + ix;
#pragma unroll
                for (int b = 0; b < NBA; ++b) {
                    const T v = src[b][gi];
#pragma unroll
                    for (int j = 0; j < MAXJ; ++j) acc[j][b] += beta[j] * v;
                }
            }
        }
    }
    // block reduction in FP64
    const int wave = tid >> 6;
#pragma unroll
    for (int j = 0; j < MAXJ; ++j)
#pragma unroll
        for (int b = 0; b < NBA; ++b) {
            double v = (double)acc[j][b];
#pragma unroll
            for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
            if ((tid & 63) == 0) red[wave][j * NBA + b] = v;
        }
    __syncthreads();
    if (tid < MAXJ * NBA) {
        const int j = tid / NBA, b = tid % NBA;
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) s += red[w][tid];
        const int cbase = a.atom_coef_off[atom];
        const int count = a.atom_coef_off[atom + 1] - cbase;
        if (b0 + b < a.nbands && j < count) a.coef[(int64_t)(b0 + b) * a.ncoef + cbase + j] = s * a.dv;
    }
}

// ------------------------------------------------------------------ host launchers ----
template <class T>
static StencilCoef<T> make_coef(const qg_grid& g, double coef) {
    const double w[4] = {8064.0, -1008.0, 128.0, -9.0};
    StencilCoef<T> s;
    const double c0 = -14350.0 / 5040.0 * coef *
                      (1.0 / (g.h[0] * g.h[0]) + 1.0 / (g.h[1] * g.h[1]) + 1.0 / (g.h[2] * g.h[2]));
    s.c0 = (T)c0;
    for (int m = 0; m < 4; ++m) {
        s.cx[m] = (T)(w[m] * coef / (5040.0 * g.h[0] * g.h[0]));
        s.cy[m] = (T)(w[m] * coef / (5040.0 * g.h[1] * g.h[1]));
        s.cz[m] = (T)(w[m] * coef / (5040.0 * g.h[2] * g.h[2]));
    }
    return s;
}

struct Plan {
    int ty, tz, ppt, nb;
    size_t lds;
};

// Tile plan: TY full rows so that a 256-thread block holds ~4-5 points per thread.
static Plan make_plan(const qg_grid& g, int dtype, int nbands, int mode) {
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    QG_REQUIRE(nx <= 128, "the fused kernel supports Nx <= 128");
    Plan p{};
    p.nb = (dtype == QG_F32) ? 2 : 1;
    if (nbands < p.nb) p.nb = 1;
    int ty = std::max(4, std::min(ny, 1280 / nx));
    int ppt = (ty * nx + kBlock - 1) / kBlock;
    const int allowed[] = {1, 2, 3, 4, 5, 6, 8};
    int chosen = 8;
    for (int v : allowed)
        if (v >= ppt) { chosen = v; break; }
    if (ppt > 8) { ty = (8 * kBlock) / nx; ppt = 8; chosen = 8; }
    p.ty = ty;
    p.ppt = chosen;
    const int ntile = (ny + ty - 1) / ty;
    const int nchunk = (nbands + p.nb - 1) / p.nb;
    // slab depth: at least 8 planes, enough workgroups to fill 256 CUs several times
    int tz = nz;
    while (tz > 8 && (int64_t)ntile * nchunk * ((nz + tz - 1) / tz) < 2048) tz = (tz + 1) / 2;
    p.tz = std::max(1, tz);
    const size_t es = (dtype == QG_F32) ? 4 : 8;
    p.lds = (mode == MODE_NL) ? 0 : 2 * (size_t)p.nb * (ty + 8) * nx * es;
    QG_REQUIRE(p.lds <= 65536, "tile does not fit in LDS");
    return p;
}

template <class T, int NB, int MODE>
static void launch_fused_ppt(const Plan& p, dim3 grid, hipStream_t s, const FusedArgs& a,
                             const StencilCoef<T>& sc, const T* V, const T* psi, T* out) {
    switch (p.ppt) {
#define QG_CASE(PP) \
    case PP: hipLaunchKernelGGL((fused_hpsi_kernel<T, NB, PP, MODE>), grid, dim3(kBlock), p.lds, s, a, sc, V, psi, out); break;
        QG_CASE(1) QG_CASE(2) QG_CASE(3) QG_CASE(4) QG_CASE(5) QG_CASE(6) QG_CASE(8)
#undef QG_CASE
        default: throw Error{QG_ERR_ARG, "unsupported points-per-thread"};
    }
    QG_HIP(hipGetLastError());
}

template <class T, int MODE>
static void launch_fused(const qg_grid& g, int nbands, const T* psi, int64_t ld_in, T* out,
                         int64_t ld_out, double coef, int accumulate, const T* V,
                         qg_nonlocal* nl, hipStream_t s) {
    const int dtype = sizeof(T) == 4 ? QG_F32 : QG_F64;
    Plan p = make_plan(g, dtype, nbands, MODE);
    FusedArgs a{};
    a.nx = g.n[0]; a.ny = g.n[1]; a.nz = g.n[2];
    a.ty = p.ty; a.tz = p.tz;
    a.ntile_y = (g.n[1] + p.ty - 1) / p.ty;
    a.nslab = (g.n[2] + p.tz - 1) / p.tz;
    a.nbands = nbands;
    a.ld_in = ld_in; a.ld_out = ld_out;
    a.h0 = g.h[0]; a.h1 = g.h[1]; a.h2 = g.h[2];
    a.accumulate = accumulate;
    a.has_nl = 0;
    if (nl && nl->geo.ncoef > 0 && MODE != MODE_LAP) {
        if (nl->list_ty != p.ty) {
            std::vector<int> off, ids;
            build_tile_lists(g, nl->geo, p.ty, off, ids);
            if (ids.empty()) ids.push_back(0);
            nl->d_list_off.upload(off.data(), off.size());
            nl->d_list_ids.upload(ids.data(), ids.size());
            nl->list_ty = p.ty;
        }
        a.has_nl = 1;
        a.list_off = nl->d_list_off.p;
        a.list_ids = nl->d_list_ids.p;
        a.images = nl->d_images.p;
        a.projs = nl->d_projs.p;
        a.tables = nl->d_tables.p;
        a.intervals = nl->d_intervals.p;
        a.coef = nl->d_coef.p;
        a.ncoef = nl->geo.ncoef;
    }
    const StencilCoef<T> sc = make_coef<T>(g, coef);
    const int nchunk = (nbands + p.nb - 1) / p.nb;
    const int64_t nwg = (int64_t)a.ntile_y * a.nslab * nchunk;
    QG_REQUIRE(nwg < (1LL << 31), "too many workgroups");
    dim3 grid((unsigned)nwg);
    if (p.nb == 2)
        launch_fused_ppt<T, 2, MODE>(p, grid, s, a, sc, V, psi, out);
    else
        launch_fused_ppt<T, 1, MODE>(p, grid, s, a, sc, V, psi, out);
}

template <class T>
static void launch_project(qg_nonlocal* nl, int nbands, const T* psi, int64_t ld, double* coef,
                           hipStream_t s) {
    const qg_grid& g = nl->gcopy;
    ProjArgs a{};
    a.nx = g.n[0]; a.ny = g.n[1]; a.nz = g.n[2];
    a.natoms = nl->geo.natoms;
    a.nbands = nbands;
    a.ld = ld;
    a.h0 = g.h[0]; a.h1 = g.h[1]; a.h2 = g.h[2];
    a.dv = g.dvol();
    a.atom_img_off = nl->d_atom_img_off.p;
    a.atom_coef_off = nl->d_atom_coef_off.p;
    a.images = nl->d_images.p;
    a.projs = nl->d_projs.p;
    a.tables = nl->d_tables.p;
    a.intervals = nl->d_intervals.p;
    a.coef = coef;
    a.ncoef = nl->geo.ncoef;
    constexpr int NBA = 4;
    const int nchunk = (nbands + NBA - 1) / NBA;
    dim3 grid((unsigned)((int64_t)a.natoms * nchunk));
    QG_REQUIRE(nl->geo.max_coef_per_atom <= 8, "at most 8 projector coefficients per atom");
    if (nl->geo.max_coef_per_atom <= 4)
        hipLaunchKernelGGL((project_kernel<T, NBA, 4>), grid, dim3(kBlock), 0, s, a, psi);
    else
        hipLaunchKernelGGL((project_kernel<T, NBA, 8>), grid, dim3(kBlock), 0, s, a, psi);
    QG_HIP(hipGetLastError());
}

// Public-to-internal helpers used by the LOBPCG driver.
void hamiltonian_apply(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi,
                       int64_t ld_out, hipStream_t s) {
    const qg_grid& g = h->grid;
    qg_nonlocal* nl = (h->nl && h->nl->geo.ncoef > 0) ? h->nl : nullptr;
    if (h->timing) QG_HIP(hipEventRecord(h->ev[0], s));
    if (nl) {
        nl->d_coef.alloc((size_t)nbands * nl->geo.ncoef);
        if (dtype == QG_F64)
            launch_project<double>(nl, nbands, (const double*)psi, ld, nl->d_coef.p, s);
        else
            launch_project<float>(nl, nbands, (const float*)psi, ld, nl->d_coef.p, s);
    }
    if (h->timing) QG_HIP(hipEventRecord(h->ev[1], s));
    if (dtype == QG_F64)
        launch_fused<double, MODE_HAM>(g, nbands, (const double*)psi, ld, (double*)hpsi, ld_out, -0.5, 0,
                                       h->V64.p, nl, s);
    else
        launch_fused<float, MODE_HAM>(g, nbands, (const float*)psi, ld, (float*)hpsi, ld_out, -0.5, 0,
                                      h->V32.p, nl, s);
    if (h->timing) {
        QG_HIP(hipEventRecord(h->ev[2], s));
        h->timed_valid = true;
    }
}

}  // namespace qg

using namespace qg;

static void check_block(const qg_grid* g, int dtype, int nbands, const void* psi, int64_t ld,
                        const void* out, int64_t ld_out) {
    QG_REQUIRE(g, "null grid");
    QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "dtype must be QG_F64 or QG_F32");
    QG_REQUIRE(nbands >= 0, "negative band count");
    QG_REQUIRE(nbands == 0 || (psi && out), "null psi/out");
    QG_REQUIRE(ld >= g->ng() && ld_out >= g->ng(), "leading dimension smaller than the grid");
    QG_REQUIRE(psi != out, "in-place application is not supported (psi must differ from out)");
}

extern "C" {

int qg_laplacian8_apply(const qg_grid* g, int dtype, int nbands, const void* psi, int64_t ld_psi,
                        void* out, int64_t ld_out, double coef, int accumulate, void* stream) {
    return guarded([&] {
        check_block(g, dtype, nbands, psi, ld_psi, out, ld_out);
        if (nbands == 0) return;
        hipStream_t s = as_stream(stream);
        if (dtype == QG_F64)
            launch_fused<double, MODE_LAP>(*g, nbands, (const double*)psi, ld_psi, (double*)out, ld_out, coef,
                                           accumulate, nullptr, nullptr, s);
        else
            launch_fused<float, MODE_LAP>(*g, nbands, (const float*)psi, ld_psi, (float*)out, ld_out, coef,
                                          accumulate, nullptr, nullptr, s);
    });
}

int qg_nonlocal_create(const qg_grid* g, int natoms, const double* xyz, const int* species,
                       int nspecies, const int* nproj, const qg_proj_spec* specs, qg_nonlocal** out) {
    return guarded([&] {
        QG_REQUIRE(g && out, "null argument");
        QG_REQUIRE(natoms >= 0 && nspecies >= 0, "negative counts");
        QG_REQUIRE(natoms == 0 || (xyz && species && nproj && (specs || nspecies == 0)), "null arrays");
        qg_nonlocal* nl = new qg_nonlocal;
        try {
            nl->gcopy = *g;
            nl->grid = &nl->gcopy;
            build_nl_geometry(*g, natoms, xyz, species, nspecies, nproj, specs, nl->geo);
            const auto& geo = nl->geo;
            if (!geo.projs.empty()) nl->d_projs.upload(geo.projs.data(), geo.projs.size());
            if (!geo.tables.empty()) nl->d_tables.upload(geo.tables.data(), geo.tables.size());
            if (!geo.images.empty()) nl->d_images.upload(geo.images.data(), geo.images.size());
            if (!geo.intervals.empty()) nl->d_intervals.upload(geo.intervals.data(), geo.intervals.size());
            nl->d_atom_img_off.upload(geo.atom_img_off.data(), geo.atom_img_off.size());
            nl->d_atom_coef_off.upload(geo.atom_coef_off.data(), geo.atom_coef_off.size());
        } catch (...) {
            delete nl;
            throw;
        }
        *out = nl;
    });
}

int qg_nonlocal_destroy(qg_nonlocal* nl) {
    delete nl;
    return QG_OK;
}

int qg_nonlocal_num_coef(const qg_nonlocal* nl) { return nl ? nl->geo.ncoef : -1; }

int qg_nonlocal_project(qg_nonlocal* nl, int dtype, int nbands, const void* psi, int64_t ld,
                        double* coef, void* stream) {
    return guarded([&] {
        QG_REQUIRE(nl, "null nonlocal handle");
        QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "bad dtype");
        QG_REQUIRE(ld >= nl->gcopy.ng(), "leading dimension smaller than the grid");
        if (nbands == 0 || nl->geo.ncoef == 0) return;
        QG_REQUIRE(psi && coef, "null pointer");
        hipStream_t s = as_stream(stream);
        if (dtype == QG_F64)
            launch_project<double>(nl, nbands, (const double*)psi, ld, coef, s);
        else
            launch_project<float>(nl, nbands, (const float*)psi, ld, coef, s);
    });
}

int qg_nonlocal_apply(qg_nonlocal* nl, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi,
                      int64_t ld_out, int accumulate, void* stream) {
    return guarded([&] {
        QG_REQUIRE(nl, "null nonlocal handle");
        check_block(&nl->gcopy, dtype, nbands, psi, ld, hpsi, ld_out);
        if (nbands == 0) return;
        hipStream_t s = as_stream(stream);
        if (nl->geo.ncoef == 0) {
            // no projectors: out = accumulate ? out : 0, through the same kernel
            if (dtype == QG_F64)
                launch_fused<double, MODE_NL>(nl->gcopy, nbands, (const double*)psi, ld, (double*)hpsi, ld_out,
                                              0.0, accumulate, nullptr, nullptr, s);
            else
                launch_fused<float, MODE_NL>(nl->gcopy, nbands, (const float*)psi, ld, (float*)hpsi, ld_out, 0.0,
                                             accumulate, nullptr, nullptr, s);
            return;
        }
        nl->d_coef.alloc((size_t)nbands * nl->geo.ncoef);
        if (dtype == QG_F64) {
            launch_project<double>(nl, nbands, (const double*)psi, ld, nl->d_coef.p, s);
            launch_fused<double, MODE_NL>(nl->gcopy, nbands, (const double*)psi, ld, (double*)hpsi, ld_out, 0.0,
                                          accumulate, nullptr, nl, s);
        } else {
            launch_project<float>(nl, nbands, (const float*)psi, ld, nl->d_coef.p, s);
            launch_fused<float, MODE_NL>(nl->gcopy, nbands, (const float*)psi, ld, (float*)hpsi, ld_out, 0.0,
                                         accumulate, nullptr, nl, s);
        }
    });
}

int qg_hamiltonian_create(const qg_grid* g, const double* V, qg_nonlocal* nl, qg_ham** out) {
    return guarded([&] {
        QG_REQUIRE(g && V && out, "null argument");
        if (nl)
            QG_REQUIRE(nl->gcopy.n[0] == g->n[0] && nl->gcopy.n[1] == g->n[1] && nl->gcopy.n[2] == g->n[2],
                       "nonlocal handle was built for another grid");
        qg_ham* h = new qg_ham;
        h->grid = *g;
        h->nl = nl;
        try {
            const int64_t ng = g->ng();
            h->V64.upload(V, ng);
            std::vector<float> v32(V, V + ng);
            h->V32.upload(v32.data(), ng);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int qg_hamiltonian_destroy(qg_ham* h) {
    delete h;
    return QG_OK;
}

int qg_hamiltonian_set_potential(qg_ham* h, const double* V) {
    return guarded([&] {
        QG_REQUIRE(h && V, "null argument");
        const int64_t ng = h->grid.ng();
        h->V64.upload(V, ng);
        std::vector<float> v32(V, V + ng);
        h->V32.upload(v32.data(), ng);
    });
}

int qg_hamiltonian_apply(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi,
                         int64_t ld_out, void* stream) {
    return guarded([&] {
        QG_REQUIRE(h, "null hamiltonian");
        check_block(&h->grid, dtype, nbands, psi, ld, hpsi, ld_out);
        if (nbands == 0) return;
        hamiltonian_apply(h, dtype, nbands, psi, ld, hpsi, ld_out, as_stream(stream));
    });
}

int qg_hamiltonian_set_timing(qg_ham* h, int enable) {
    return guarded([&] {
        QG_REQUIRE(h, "null hamiltonian");
        if (enable)
            for (auto& e : h->ev)
                if (!e) QG_HIP(hipEventCreate(&e));
        h->timing = enable != 0;
        h->timed_valid = false;
    });
}

int qg_hamiltonian_last_times(qg_ham* h, double* t_project_ms, double* t_fused_ms) {
    return guarded([&] {
        QG_REQUIRE(h && t_project_ms && t_fused_ms, "null argument");
        QG_REQUIRE(h->timing && h->timed_valid, "no timed apply recorded (enable timing first)");
        QG_HIP(hipEventSynchronize(h->ev[2]));
        float a = 0.f, b = 0.f;
        QG_HIP(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
        QG_HIP(hipEventElapsedTime(&b, h->ev[1], h->ev[2]));
        *t_project_ms = a;
        *t_fused_ms = b;
    });
}

}  // extern "C"
