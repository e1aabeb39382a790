// oracle.cpp — CPU FP64 restatement of QUMASUN's Kohn-Sham H·psi + block-LOBPCG path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle for the MI355X engine: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product
// library (qumasun_amd/) never links, calls or falls back to it.
//
// Every function restates the reference algorithm named in its comment (file:line into
// QUMASUN's source/), in the same arithmetic order where that order is visible, written
// from scratch (no reference source is copied).  Pinning (what checks this file) is listed
// per function; see tests/test_oracle_*.py.
//
// Dense linear algebra (the reference's cblas_dgemm / LAPACK DSYGVX, whose MKL/NEC builds
// are not in the reference tree) is delegated to the reference LAPACK routines shipped in
// scipy's OpenBLAS (scipy_dgemm_, scipy_dsygvx_), called with the reference's arguments.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#include <algorithm>
#include "../include/qumasun_gpu.h"   // qg_proj_spec (plain C struct) only

extern "C" {
void scipy_dgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k,
                  const double* alpha, const double* a, const int* lda, const double* b,
                  const int* ldb, const double* beta, double* c, const int* ldc);
void scipy_dsygvx_(const int* itype, const char* jobz, const char* range, const char* uplo,
                   const int* n, double* a, const int* lda, double* b, const int* ldb,
                   const double* vl, const double* vu, const int* il, const int* iu,
                   const double* abstol, int* m, double* w, double* z, const int* ldz,
                   double* work, const int* lwork, int* iwork, int* ifail, int* info);
void scipy_openblas_set_num_threads(int);
}

namespace {

// column-major C = alpha*op(A)*op(B) + beta*C  (cblas_dgemm(ColMajor, ...))
void gemm(bool ta, bool tb, int m, int n, int k, double alpha, const double* a, int lda,
          const double* b, int ldb, double beta, double* c, int ldc) {
    if (m <= 0 || n <= 0) return;
    const char cta = ta ? 'T' : 'N', ctb = tb ? 'T' : 'N';
    scipy_dgemm_(&cta, &ctb, &m, &n, &k, &alpha, a, &lda, b, &ldb, &beta, c, &ldc);
}

}  // namespace

#define OR_API extern "C"

OR_API void or_set_threads(int n) { scipy_openblas_set_num_threads(n); }

// ---------------------------------------------------------------------------------------
// 8th-order periodic Laplacian: out += coef * Lap8 p.
// Restates PasteHalo_f4 + Laplasian8th_halo_f + Laplasian8th_halo + Laplasian8th
// (GridDifference8th.h:20-79, :83-149, :153-170, :178-185): the field is first copied
// into a buffer with a 4-point periodic halo on the six faces, then the 25-point star is
// evaluated as c0*p + x-part + y-part + z-part and added to out.
// Pinned by: oracle/_ref (the reference header itself), plane-wave symbol, brute force.
OR_API void or_laplacian8(int nx, int ny, int nz, double dx, double dy, double dz, double coef,
                   const double* p, double* out) {
    const int W = 4;
    const int sx = nx + 2 * W, sy = ny + 2 * W, sz = nz + 2 * W;
    std::vector<double> h((size_t)sx * sy * sz, 0.0);
    auto H = [&](int ix, int iy, int iz) -> double& {
        return h[(size_t)(ix + W) + (size_t)sx * ((iy + W) + (size_t)sy * (iz + W))];
    };
    auto P = [&](int ix, int iy, int iz) {
        ix = (ix % nx + nx) % nx; iy = (iy % ny + ny) % ny; iz = (iz % nz + nz) % nz;
        return p[ix + (size_t)nx * (iy + (size_t)ny * iz)];
    };
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) H(ix, iy, iz) = P(ix, iy, iz);
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int m = 1; m <= W; ++m) { H(-m, iy, iz) = P(-m, iy, iz); H(nx - 1 + m, iy, iz) = P(nx - 1 + m, iy, iz); }
    for (int iz = 0; iz < nz; ++iz)
        for (int ix = 0; ix < nx; ++ix)
            for (int m = 1; m <= W; ++m) { H(ix, -m, iz) = P(ix, -m, iz); H(ix, ny - 1 + m, iz) = P(ix, ny - 1 + m, iz); }
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix)
            for (int m = 1; m <= W; ++m) { H(ix, iy, -m) = P(ix, iy, -m); H(ix, iy, nz - 1 + m) = P(ix, iy, nz - 1 + m); }

    // Fornberg's 8th-order central weights times 5040: -14350, 8064, -1008, 128, -9
    const double c0 = -14350.0 / 5040.0 * coef * (1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz));
    const double w[5] = {0.0, 8064.0, -1008.0, 128.0, -9.0};
    double cx[5], cy[5], cz[5];
    for (int m = 1; m <= 4; ++m) {
        cx[m] = w[m] * coef / (5040.0 * dx * dx);
        cy[m] = w[m] * coef / (5040.0 * dy * dy);
        cz[m] = w[m] * coef / (5040.0 * dz * dz);
    }
    const int64_t s1 = 1, s2 = sx, s3 = (int64_t)sx * sy;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const double* q = &H(ix, iy, iz);
                double d = c0 * q[0];
                d += cx[4] * (q[4 * s1] + q[-4 * s1]) + cx[3] * (q[3 * s1] + q[-3 * s1]) +
                     cx[2] * (q[2 * s1] + q[-2 * s1]) + cx[1] * (q[s1] + q[-s1]);
                d += cy[4] * (q[4 * s2] + q[-4 * s2]) + cy[3] * (q[3 * s2] + q[-3 * s2]) +
                     cy[2] * (q[2 * s2] + q[-2 * s2]) + cy[1] * (q[s2] + q[-s2]);
                d += cz[4] * (q[4 * s3] + q[-4 * s3]) + cz[3] * (q[3 * s3] + q[-3 * s3]) +
                     cz[2] * (q[2 * s3] + q[-2 * s3]) + cz[1] * (q[s3] + q[-s3]);
                out[ix + (size_t)nx * (iy + (size_t)ny * iz)] += d;
            }
}

// ---------------------------------------------------------------------------------------
// Radial interpolation on the log grid: RadialGrid2::GetValueByXi (RadialGrid2.h:81-136):
// linear in the first and last intervals, cubic Hermite with 2nd-order slopes in the
// second and second-to-last, quintic Hermite with 4th-order first/second derivatives
// inside, zero beyond the table.
// Pinned by: polynomial exactness per branch (tests/test_oracle_radial.py).
OR_API double or_radial_by_xi(double xi, const double* v, int n, double xi_min, double xi_delta) {
    const double rel = xi - xi_min;
    const double fl = std::floor(rel / xi_delta);
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
        // quintic p(t) = f0 + g0 t + h0/2 t^2 + a3 t^3 + a4 t^4 + a5 t^5 matching (f,g,h) at 1
        const double e0 = f1 - (f0 + g0 + h0 / 2.0);
        const double e1 = g1 - (g0 + h0);
        const double e2 = h1 - h0;
        const double a3 = 10.0 * e0 - 4.0 * e1 + 0.5 * e2;
        const double a4 = -15.0 * e0 + 7.0 * e1 - e2;
        const double a5 = f1 - f0 - g0 - h0 / 2.0 - a3 - a4;
        return f0 + t * (g0 + t * (h0 / 2.0 + t * (a3 + t * (a4 + t * a5))));
    }
    // i == 1 or i == n-3: cubic Hermite with centred-difference slopes
    const double v0 = v[i], v1 = v[i + 1];
    const double s1 = (v[i + 2] - v0) / 2.0, s0 = (v1 - v[i - 1]) / 2.0;
    const double a = s1 + s0 - 2.0 * (v1 - v0);
    const double b = -s1 - 2.0 * s0 + 3.0 * (v1 - v0);
    return a * t * t * t + b * t * t + s0 * t + v0;
}

// ---------------------------------------------------------------------------------------
// Nonlocal KB projectors on one process that holds the whole periodic grid.
struct OrBlock { int b[3], e[3], shift[3]; };
struct OrAtom {
    double R[3];
    int max_l;
    std::vector<int> proj;                      // indices into or_nl::specs
    std::vector<std::vector<OrBlock>> blocks;   // per l
    std::vector<int> size;                      // per l: total points of its blocks
    std::vector<std::vector<double>> radial;    // per projector k: p(r) on its l-blocks
    std::vector<std::vector<double>> ylm;       // (max_l+1)^2 tables, index l*l+m+l
};
struct or_nl {
    int n[3];
    double h[3];
    std::vector<qg_proj_spec> specs;
    std::vector<std::vector<double>> tables;
    std::vector<OrAtom> atoms;
    int ncoef;
};

static inline double Y00c() { return 1.0 / std::sqrt(4.0 * M_PI); }

// GridForSphere (GridSubgrid.h:483-492) + GetOverlapPeriodic (GridSubgrid.h:381-430) with
// grid1 = the whole grid: the sphere's box split into its periodic-image blocks, nz-major.
static std::vector<OrBlock> sphere_blocks(const int n[3], const double h[3], const double R[3], double rc) {
    int bb[3], be[3];
    for (int d = 0; d < 3; ++d) {
        bb[d] = (int)std::ceil((R[d] - rc) / h[d] - 0.5);
        be[d] = (int)std::floor((R[d] + rc) / h[d] + 0.5) + 1;
    }
    auto first_gt = [](int x, int y) { return (x >= 0) ? (x / y) + 1 : ((x + 1) / y); };
    int ns[3], ne[3];
    for (int d = 0; d < 3; ++d) {
        ns[d] = first_gt(0 - be[d], n[d]);
        ne[d] = -first_gt(bb[d] - n[d], n[d]);
    }
    std::vector<OrBlock> out;
    for (int kz = ns[2]; kz <= ne[2]; ++kz)
        for (int ky = ns[1]; ky <= ne[1]; ++ky)
            for (int kx = ns[0]; kx <= ne[0]; ++kx) {
                const int k[3] = {kx, ky, kz};
                OrBlock blk;
                for (int d = 0; d < 3; ++d) {
                    blk.b[d] = std::max(0, bb[d] + k[d] * n[d]);
                    blk.e[d] = std::min(n[d], be[d] + k[d] * n[d]);
                    blk.shift[d] = k[d] * n[d];
                }
                out.push_back(blk);
            }
    return out;
}

static int64_t block_size(const OrBlock& b) {
    return (int64_t)(b.e[0] - b.b[0]) * (b.e[1] - b.b[1]) * (b.e[2] - b.b[2]);
}

// real Y_lm of RealSphericalHarmonics.h:149-161 for l <= 1 (m = -1: y, 0: z, +1: x)
static double ylm_eval(int l, int m, double x, double y, double z) {
    if (l == 0) return std::sqrt(1.0 / (4.0 * M_PI));
    const double c = (m != 0) ? std::sqrt(3.0 * 1.0 / (2.0 * 2.0 * M_PI)) : std::sqrt(3.0 / (4.0 * M_PI));
    if (m == -1) return c * 1.0 * y;
    if (m == 0) return c * z * 1.0;
    return c * 1.0 * x;
}

// Visit every point of an atom's l-blocks in table order with its offset from the nucleus
// (the loop of SetBlockYlmAroundAny, SetBlockYlm.h:10-53, and of
// mCreateProjectorRadialOnBlock, PseudoPotOperator_mpi_v2.h:128-176).
template <class F>
static void for_block_points(const std::vector<OrBlock>& blocks, const double h[3], const double R[3], F&& f) {
    int64_t off = 0;
    for (const auto& b : blocks) {
        const int sxb = b.e[0] - b.b[0], syb = b.e[1] - b.b[1];
        for (int iz = b.b[2]; iz < b.e[2]; ++iz) {
            const double z = h[2] * (double)(iz - b.shift[2]) - R[2];
            const double zz = z * z;
            for (int iy = b.b[1]; iy < b.e[1]; ++iy) {
                const double y = h[1] * (double)(iy - b.shift[1]) - R[1];
                const double yy_zz = y * y + zz;
                for (int ix = b.b[0]; ix < b.e[0]; ++ix) {
                    const double x = h[0] * (double)(ix - b.shift[0]) - R[0];
                    const double rr = x * x + yy_zz;
                    const int64_t i = off + (ix - b.b[0]) + sxb * ((iy - b.b[1]) + (int64_t)syb * (iz - b.b[2]));
                    f(i, ix, iy, iz, x, y, z, rr);
                }
            }
        }
        off += block_size(b);
    }
}

static double radial_value(const qg_proj_spec& s, const std::vector<double>& tab, double rr) {
    if (s.radial_kind == QG_RADIAL_GAUSSIAN) {
        if (s.cutoff * s.cutoff < rr) return 0.0;
        return s.amplitude * std::exp(-rr / (2.0 * s.sigma * s.sigma));
    }
    // PseudoPotOperator_mpi_v2.h:154-166 with the j-averaged table
    if (s.r_min * s.r_min > rr) return tab[0];
    if (s.cutoff * s.cutoff < rr) return 0.0;
    return or_radial_by_xi(0.5 * std::log(rr), tab.data(), s.n_table, s.xi_min, s.xi_delta);
}

// UpdatePosition (PseudoPotOperator_mpi_v2.h:231-248): blocks per (atom, l)
// (PseudoPotOperator_vlocal.h:142-156), Ylm tables (tCreateProjectorYlmOnBlock,
// PseudoPotOperator_vlocal.h:872-897) and radial tables (mCreateProjectorRadialOnBlock).
OR_API or_nl* or_nl_create(int nx, int ny, int nz, double dx, double dy, double dz, int natoms,
                    const double* xyz, const int* species, int nspecies, const int* nproj,
                    const qg_proj_spec* specs) {
    or_nl* nl = new or_nl;
    nl->n[0] = nx; nl->n[1] = ny; nl->n[2] = nz;
    nl->h[0] = dx; nl->h[1] = dy; nl->h[2] = dz;
    std::vector<int> first(nspecies + 1, 0);
    for (int s = 0; s < nspecies; ++s) first[s + 1] = first[s] + nproj[s];
    for (int k = 0; k < first[nspecies]; ++k) {
        nl->specs.push_back(specs[k]);
        if (specs[k].radial_kind == QG_RADIAL_TABLE)
            nl->tables.emplace_back(specs[k].table, specs[k].table + specs[k].n_table);
        else
            nl->tables.emplace_back();
    }
    nl->ncoef = 0;
    for (int a = 0; a < natoms; ++a) {
        OrAtom at;
        for (int d = 0; d < 3; ++d) at.R[d] = xyz[3 * a + d];
        const int s = species[a];
        at.max_l = 0;
        for (int k = first[s]; k < first[s + 1]; ++k) {
            at.proj.push_back(k);
            at.max_l = std::max(at.max_l, specs[k].l);
            nl->ncoef += 2 * specs[k].l + 1;
        }
        at.blocks.resize(at.max_l + 1);
        at.size.assign(at.max_l + 1, 0);
        for (int l = 0; l <= at.max_l; ++l) {
            // cutoff_r[l]: the cutoff of the species' projector(s) with this l
            double rc = 0.0;
            for (int k : at.proj) if (specs[k].l == l) rc = specs[k].cutoff;
            at.blocks[l] = sphere_blocks(nl->n, nl->h, at.R, rc);
            int64_t tot = 0;
            for (auto& b : at.blocks[l]) tot += block_size(b);
            at.size[l] = (int)tot;
        }
        at.ylm.resize((at.max_l + 1) * (at.max_l + 1));
        for (int l = 0; l <= at.max_l; ++l)
            for (int m = -l; m <= l; ++m) {
                auto& t = at.ylm[l * l + m + l];
                t.assign(at.size[l], 0.0);
                const double org = (l == 0) ? 1.0 / std::sqrt(4.0 * M_PI) : 0.0;
                for_block_points(at.blocks[l], nl->h, at.R,
                                 [&](int64_t i, int, int, int, double x, double y, double z, double rr) {
                                     if (1.0e-14 > rr) { t[i] = org; return; }
                                     const double r = std::sqrt(rr);
                                     t[i] = ylm_eval(l, m, x / r, y / r, z / r);
                                 });
            }
        for (int k : at.proj) {
            const int l = specs[k].l;
            std::vector<double> t(at.size[l], 0.0);
            for_block_points(at.blocks[l], nl->h, at.R,
                             [&](int64_t i, int, int, int, double, double, double, double rr) {
                                 t[i] = radial_value(nl->specs[k], nl->tables[k], rr);
                             });
            at.radial.push_back(std::move(t));
        }
        nl->atoms.push_back(std::move(at));
    }
    return nl;
}

OR_API void or_nl_destroy(or_nl* nl) { delete nl; }
OR_API int or_nl_num_coef(const or_nl* nl) { return nl->ncoef; }

// Total number of block points of atom a at angular momentum l (the table length).
OR_API int or_nl_block_points(const or_nl* nl, int a, int l) {
    if (a < 0 || a >= (int)nl->atoms.size() || l > nl->atoms[a].max_l) return -1;
    return nl->atoms[a].size[l];
}

// Dense projector column j (length Ng) in coefficient order: beta_j at every grid point,
// summed over periodic images.  For brute-force checks only.
OR_API void or_nl_dense_beta(const or_nl* nl, int j, double* col) {
    const int64_t ng = (int64_t)nl->n[0] * nl->n[1] * nl->n[2];
    std::fill(col, col + ng, 0.0);
    int idx = 0;
    for (const auto& at : nl->atoms) {
        for (size_t kk = 0; kk < at.proj.size(); ++kk) {
            const int l = nl->specs[at.proj[kk]].l;
            for (int m1 = 0; m1 < 2 * l + 1; ++m1, ++idx) {
                if (idx != j) continue;
                const auto& rad = at.radial[kk];
                for_block_points(at.blocks[l], nl->h, at.R,
                                 [&](int64_t i, int ix, int iy, int iz, double, double, double, double) {
                                     const double yv = (l == 0) ? Y00c() : at.ylm[l * l + m1][i];
                                     col[ix + (int64_t)nl->n[0] * (iy + (int64_t)nl->n[1] * iz)] += rad[i] * yv;
                                 });
            }
        }
    }
}

// Dense k-point projector column j: beta_j(x_u - R) exp(i k . x_u) summed over the periodic
// images, x_u = h (i - shift) the image-frame coordinate of grid point i (the projector blocks of
// mCreateProjectorRadialOnBlock, PseudoPotOperator_kpoint_v2.h:287-420, whose phase is
// cos/sin(k . x_u) with k . x_u = gx_dx (ix - shift_x) + ...).  For brute-force checks only.
OR_API void or_nl_dense_beta_k(const or_nl* nl, int j, double kx, double ky, double kz, double* re,
                               double* im) {
    const int64_t ng = (int64_t)nl->n[0] * nl->n[1] * nl->n[2];
    std::fill(re, re + ng, 0.0);
    std::fill(im, im + ng, 0.0);
    int idx = 0;
    for (const auto& at : nl->atoms) {
        for (size_t kk = 0; kk < at.proj.size(); ++kk) {
            const int l = nl->specs[at.proj[kk]].l;
            for (int m1 = 0; m1 < 2 * l + 1; ++m1, ++idx) {
                if (idx != j) continue;
                const auto& rad = at.radial[kk];
                int64_t off = 0;
                for (const auto& b : at.blocks[l]) {
                    const int sxb = b.e[0] - b.b[0], syb = b.e[1] - b.b[1];
                    for (int iz = b.b[2]; iz < b.e[2]; ++iz)
                        for (int iy = b.b[1]; iy < b.e[1]; ++iy)
                            for (int ix = b.b[0]; ix < b.e[0]; ++ix) {
                                const int64_t i = off + (ix - b.b[0]) + sxb * ((iy - b.b[1]) + (int64_t)syb * (iz - b.b[2]));
                                const double yv = (l == 0) ? Y00c() : at.ylm[l * l + m1][i];
                                const double th = kx * nl->h[0] * (double)(ix - b.shift[0]) +
                                                  ky * nl->h[1] * (double)(iy - b.shift[1]) +
                                                  kz * nl->h[2] * (double)(iz - b.shift[2]);
                                const int64_t g = ix + (int64_t)nl->n[0] * (iy + (int64_t)nl->n[1] * iz);
                                re[g] += rad[i] * yv * std::cos(th);
                                im[g] += rad[i] * yv * std::sin(th);
                            }
                    off += block_size(b);
                }
            }
        }
    }
}

// Per-projector KB energies in coefficient order.
OR_API void or_nl_energies(const or_nl* nl, double* e) {
    int idx = 0;
    for (const auto& at : nl->atoms)
        for (int k : at.proj)
            for (int m1 = 0; m1 < 2 * nl->specs[k].l + 1; ++m1) e[idx++] = nl->specs[k].energy;
}

// mInnerNonlocalPsi_v2 (PseudoPotOperator_mpi_v2.h:336-463) on one process:
// cut psi onto each (atom, l) block table (CutSubgridByRanges, GridSubgrid.h:435-441),
// then coef = dV*Y00*sum(p*psi) for l = 0 and dV*sum(p*Ylm*psi) for l = 1.
OR_API void or_nl_project(const or_nl* nl, const double* psi, double* coef) {
    const double dV = nl->h[0] * nl->h[1] * nl->h[2];
    const double Y_00 = 1.0 / std::sqrt(4.0 * M_PI);
    int idx = 0;
    std::vector<double> cut;
    for (const auto& at : nl->atoms) {
        std::vector<std::vector<double>> cuts(at.max_l + 1);
        for (int l = 0; l <= at.max_l; ++l) {
            cuts[l].assign(at.size[l], 0.0);
            for_block_points(at.blocks[l], nl->h, at.R,
                             [&](int64_t i, int ix, int iy, int iz, double, double, double, double) {
                                 cuts[l][i] = psi[ix + (int64_t)nl->n[0] * (iy + (int64_t)nl->n[1] * iz)];
                             });
        }
        for (size_t kk = 0; kk < at.proj.size(); ++kk) {
            const int l = nl->specs[at.proj[kk]].l;
            const auto& proj = at.radial[kk];
            const auto& ps = cuts[l];
            if (l == 0) {
                double inner = 0.0;
                for (int i = 0; i < at.size[l]; ++i) inner += proj[i] * ps[i];
                coef[idx++] = inner * dV * Y_00;
            } else {
                for (int m1 = 0; m1 < 2 * l + 1; ++m1) {
                    const auto& Y = at.ylm[l * l + m1];
                    double inner = 0.0;
                    for (int i = 0; i < at.size[l]; ++i) inner += proj[i] * Y[i] * ps[i];
                    coef[idx++] = inner * dV;
                }
            }
        }
    }
}

// The apply half of ProjectionPP (PseudoPotOperator_mpi_v2.h:506-566) given coefficients:
// per (atom, l) accumulate E*coef*p*Ylm (l = 0: E*Y00*coef*p) on the block table, then
// scatter-add it into Hp (AddSubgridByRanges, GridSubgrid.h:444-450).
OR_API void or_nl_apply_coef(const or_nl* nl, const double* coef, double* hp) {
    const double Y_00 = 1.0 / std::sqrt(4.0 * M_PI);
    int idx = 0;
    for (const auto& at : nl->atoms) {
        std::vector<std::vector<double>> cut(at.max_l + 1);
        for (int l = 0; l <= at.max_l; ++l) cut[l].assign(at.size[l], 0.0);
        for (size_t kk = 0; kk < at.proj.size(); ++kk) {
            const int l = nl->specs[at.proj[kk]].l;
            const double E = nl->specs[at.proj[kk]].energy;
            const auto& proj = at.radial[kk];
            if (l == 0) {
                double inner = coef[idx++];
                inner *= E;
                inner *= Y_00;
                for (int i = 0; i < at.size[l]; ++i) cut[l][i] += proj[i] * inner;
            } else {
                for (int m1 = 0; m1 < 2 * l + 1; ++m1) {
                    const auto& Y = at.ylm[l * l + m1];
                    double inner = coef[idx++];
                    inner *= E;
                    for (int i = 0; i < at.size[l]; ++i) cut[l][i] += proj[i] * Y[i] * inner;
                }
            }
        }
        for (int l = 0; l <= at.max_l; ++l)
            for_block_points(at.blocks[l], nl->h, at.R,
                             [&](int64_t i, int ix, int iy, int iz, double, double, double, double) {
                                 hp[ix + (int64_t)nl->n[0] * (iy + (int64_t)nl->n[1] * iz)] += cut[l][i];
                             });
    }
}

// ProjectionPP (PseudoPotOperator_mpi_v2.h:465-570): Hp += V_NL psi for one band.
OR_API void or_nl_apply(const or_nl* nl, const double* psi, double* hp) {
    std::vector<double> coef(nl->ncoef);
    or_nl_project(nl, psi, coef.data());
    or_nl_apply_coef(nl, coef.data(), hp);
}

// mHamiltonianMatrix_ddm (qumasun_mpi.matrix.h:21-43): Hp = 0; Hp = V*p
// (mPotentialMatrix_ddm :159-164); Hp += V_NL p; Hp += -1/2 Lap8 p (:126-130).
OR_API void or_hamiltonian(int nx, int ny, int nz, double dx, double dy, double dz, const double* V,
                    const or_nl* nl, const double* p, double* hp) {
    const int64_t ng = (int64_t)nx * ny * nz;
    for (int64_t i = 0; i < ng; ++i) hp[i] = 0.0;
    for (int64_t i = 0; i < ng; ++i) hp[i] = V[i] * p[i];
    if (nl) or_nl_apply(nl, p, hp);
    or_laplacian8(nx, ny, nz, dx, dy, dz, -1.0 / 2.0, p, hp);
}

// block version over nbands band-major columns
OR_API void or_hamiltonian_block(int nx, int ny, int nz, double dx, double dy, double dz, const double* V,
                          const or_nl* nl, int nbands, const double* p, double* hp) {
    const int64_t ng = (int64_t)nx * ny * nz;
    for (int b = 0; b < nbands; ++b) or_hamiltonian(nx, ny, nz, dx, dy, dz, V, nl, p + b * ng, hp + b * ng);
}

// qumasun_mpi.initialize.h:196-214: mt19937(seed), U[0,1) per point, band by band,
// Normalize(dVol) (vecmath.h:128-134), plus one extra draw per band for the guess -1/val.
OR_API void or_random_states(int64_t ng, int nbands, double dvol, uint32_t seed, double* X, double* guess) {
    std::mt19937 mt(seed);
    std::uniform_real_distribution<> dist(0.0, 1.0);
    double val = 1.0;
    for (int n = 0; n < nbands; ++n) {
        double* x = X + (int64_t)n * ng;
        for (int64_t i = 0; i < ng; ++i) x[i] = dist(mt);
        double nrm = 0.0;
        for (int64_t i = 0; i < ng; ++i) nrm += x[i] * x[i];
        const double c = 1.0 / std::sqrt(nrm * dvol);
        for (int64_t i = 0; i < ng; ++i) x[i] *= c;
        val += dist(mt);
        if (guess) guess[n] = -1.0 / val;
    }
}

// GramSchmidt (GramSchmidt.h:6-38): band n is orthogonalised against bands 0..n-1 one after
// another (each projection uses the already-updated band, dVol-weighted inner products),
// then normalised to dVol * sum psi^2 = 1.  Pinned bit for bit against the reference's own
// function (_ref) and against numpy's QR (tests/test_oracle_gram_schmidt.py).
OR_API void or_gram_schmidt(int64_t ng, int num, double* psi, double dvol) {
    for (int n = 0; n < num; ++n) {
        double* pn = psi + (int64_t)n * ng;
        for (int k = 0; k < n; ++k) {
            const double* pk = psi + (int64_t)k * ng;
            double in = 0.0;
            for (int64_t i = 0; i < ng; ++i) in += pn[i] * pk[i];
            in *= dvol;
            for (int64_t i = 0; i < ng; ++i) pn[i] -= in * pk[i];
        }
        double nrm = 0.0;
        for (int64_t i = 0; i < ng; ++i) nrm += pn[i] * pn[i];
        nrm *= dvol;
        const double inv = 1.0 / std::sqrt(nrm);
        for (int64_t i = 0; i < ng; ++i) pn[i] *= inv;
    }
}

// ---------------------------------------------------------------------------------------
// Block LOBPCG: LOBPCG::Eigen_d_multi_mpi (lobpcg_d_multi_mpi.h:884-1572) on one process
// with the BLAS Gram path (MakeMatrix_d_blas :192-250, DIRECT_S_MATRIX_XR_RP, the
// 10-iteration refresh, CHECK_NEXT_VECTOR rollback), NextVectors2 (:550-579) and the
// small-matrix recurrences NextMatrix_d (:265-460).  The MPI reductions are identities
// here, so the Compress/Expand buffer moves (:1058-1080) reduce to using the blocks in place.
// Extension (off when tol == 0): stop after an iteration whose max_k ||r_k|| <= tol.
// H·psi is supplied by a callback so tests can plug in the oracle Hamiltonian or a matrix.
typedef void (*or_opA)(void* ctx, int nbands, const double* x, double* ax);

struct OrLobpcgOut { int iterations; int rollbacks; int hpsi_calls; double max_residual; };

static void upper_dsygvx(int n, double* A, double* B, double* w, double* Z) {
    // lapack_DSYGVX (w_dsygvx.h:25-103): ITYPE 1, JOBZ V, RANGE A, UPLO U, ABSTOL 1e-13,
    // LWORK 8n; INFO > 0 is ignored as in the reference.
    const int itype = 1; const char jobz = 'V', range = 'A', uplo = 'U';
    int m = 0, info = 0; const double vl = 0, vu = 0; const int il = 1, iu = n;
    const double abstol = 1.0e-13; const int lwork = 8 * n;
    std::vector<double> work(lwork); std::vector<int> iwork(5 * n), ifail(n);
    scipy_dsygvx_(&itype, &jobz, &range, &uplo, &n, A, &n, B, &n, &vl, &vu, &il, &iu, &abstol,
                  &m, w, Z, &n, work.data(), &lwork, iwork.data(), ifail.data(), &info);
    if (info < 0) std::fprintf(stderr, "dsygvx info=%d\n", info);
}

static void copy_sub(int ncol, int nrow, double* d, int ldd, const double* s, int lds) {
    for (int j = 0; j < ncol; ++j)
        for (int i = 0; i < nrow; ++i) d[i + (int64_t)ldd * j] = s[i + (int64_t)lds * j];
}

OR_API int or_lobpcg_d_multi(int64_t ng, double dvol, int N, int iter_max, double tol,
                      double* eig, double* X, double* P, int scf_step,
                      or_opA opA, void* ctx, OrLobpcgOut* out) {
    const int64_t K = ng;
    const int n3 = 3 * N, n2 = 2 * N;
    const int64_t nn9 = (int64_t)n3 * n3;
    std::vector<double> R(K * N), AX(K * N), AR(K * N), AP(K * N), T(K * N);
    std::vector<double> Sa(nn9 * 2 + 4 * (int64_t)N * N, 0.0), Sb(nn9 * 2 + 4 * (int64_t)N * N, 0.0);
    std::vector<double> rSa(nn9), rSb(nn9), ev(n3), Z(nn9);
    const bool first_time = (scf_step == 0);
    int rollbacks = 0, hcalls = 0;
    auto xs = [&](double* base, int k) { return base + (int64_t)k * K; };

    const double sq = std::sqrt(dvol);
    for (int64_t i = 0; i < K * N; ++i) X[i] *= sq;
    opA(ctx, N, X, AX.data());
    opA(ctx, N, P, AP.data());
    hcalls += 2 * N;
    for (int k = 0; k < N; ++k) {
        double s = 0.0;
        for (int64_t i = 0; i < K; ++i) s += xs(X, k)[i] * xs(AX.data(), k)[i];
        eig[k] = s;
    }
    for (int k = 0; k < N; ++k)
        for (int64_t i = 0; i < K; ++i) xs(R.data(), k)[i] = xs(AX.data(), k)[i] - eig[k] * xs(X, k)[i];
    if (first_time) {
        std::fill(P, P + K * N, 0.0);
        std::fill(AP.begin(), AP.end(), 0.0);
    }
    if (iter_max < 1) {
        for (int64_t i = 0; i < K * N; ++i) X[i] *= 1.0 / sq;
        if (out) { out->iterations = 0; out->rollbacks = 0; out->hpsi_calls = hcalls; out->max_residual = -1; }
        return 0;
    }
    // Gram block pointers inside the 3N x 3N column-major Sa (H) / Sb (S)
    auto blk = [&](std::vector<double>& S, int bi, int bj) { return S.data() + (int64_t)N * bi + (int64_t)N * n3 * bj; };
    auto gemm_t = [&](const double* A, const double* B, double* C) {  // C = A^T B, K-length columns
        gemm(true, false, N, N, (int)K, 1.0, A, (int)K, B, (int)K, 0.0, C, n3);
    };
    bool large_diff = false;
    double maxres = -1.0;
    int iter;
    for (iter = 0; iter < iter_max; ++iter) {
        opA(ctx, N, R.data(), AR.data());
        hcalls += N;
        const bool full = (iter == 0) || (iter % 10 == 0) || large_diff;
        // MakeMatrix_d_blas: the 12 upper blocks when full, else XtHR, RtHR, RtHP, RtR.
        if (full) {
            gemm_t(X, AX.data(), blk(Sa, 0, 0));
            gemm_t(X, AP.data(), blk(Sa, 0, 2));
            gemm_t(P, AP.data(), blk(Sa, 2, 2));
            gemm_t(X, X, blk(Sb, 0, 0));
            gemm_t(X, R.data(), blk(Sb, 0, 1));
            gemm_t(X, P, blk(Sb, 0, 2));
            gemm_t(R.data(), P, blk(Sb, 1, 2));
            gemm_t(P, P, blk(Sb, 2, 2));
            large_diff = false;
        }
        gemm_t(X, AR.data(), blk(Sa, 0, 1));
        gemm_t(R.data(), AR.data(), blk(Sa, 1, 1));
        gemm_t(R.data(), AP.data(), blk(Sa, 1, 2));
        gemm_t(R.data(), R.data(), blk(Sb, 1, 1));

        const bool use_p = !(first_time && iter == 0);
        const int nb = use_p ? n3 : n2;
        // the (nb x nb) upper-triangular problem in a dense buffer
        for (int j = 0; j < nb; ++j)
            for (int i = 0; i < nb; ++i) {
                rSa[i + (int64_t)nb * j] = Sa[i + (int64_t)n3 * j];
                rSb[i + (int64_t)nb * j] = Sb[i + (int64_t)n3 * j];
            }
        upper_dsygvx(nb, rSa.data(), rSb.data(), ev.data(), Z.data());
        const int ldc = nb;
        const double* Cx = Z.data();
        const double* Cr = Z.data() + N;
        const double* Cp = Z.data() + 2 * N;
        const bool final_step = (iter + 1 == iter_max);
        // NextVectors2(x, r, p, tmp): T = R Cr (+ P Cp); P = T; T += X Cx
        gemm(false, false, (int)K, N, N, 1.0, R.data(), (int)K, Cr, ldc, 0.0, T.data(), (int)K);
        if (use_p) gemm(false, false, (int)K, N, N, 1.0, P, (int)K, Cp, ldc, 1.0, T.data(), (int)K);
        std::copy(T.begin(), T.end(), P);
        gemm(false, false, (int)K, N, N, 1.0, X, (int)K, Cx, ldc, 1.0, T.data(), (int)K);
        std::vector<double> nx(N), np(N);
        for (int j = 0; j < N; ++j) {
            double a = 0.0, b = 0.0;
            for (int64_t i = 0; i < K; ++i) { a += xs(T.data(), j)[i] * xs(T.data(), j)[i]; b += xs(P, j)[i] * xs(P, j)[i]; }
            nx[j] = a; np[j] = b;
        }
        double maxd = 0.0;
        for (int k = 0; k < N; ++k) maxd = std::max(maxd, std::fabs(1.0 - nx[k]));
        large_diff = maxd > 1.0e-3;
        if (large_diff) {
            // rollback (:1322-1346): P = R, X back to the grid normalisation, leave the loop
            std::copy(R.begin(), R.end(), P);
            for (int64_t i = 0; i < K * N; ++i) X[i] *= 1.0 / sq;
            ++rollbacks;
            ++iter;
            break;
        }
        std::copy(T.begin(), T.end(), X);
        for (int i = 0; i < N; ++i) eig[i] = ev[i];
        bool stop = final_step;
        if (!final_step) {
            // NextVectors2(Ax, Ar, Ap, tmp); AX = T
            gemm(false, false, (int)K, N, N, 1.0, AR.data(), (int)K, Cr, ldc, 0.0, T.data(), (int)K);
            if (use_p) gemm(false, false, (int)K, N, N, 1.0, AP.data(), (int)K, Cp, ldc, 1.0, T.data(), (int)K);
            std::copy(T.begin(), T.end(), AP.begin());
            gemm(false, false, (int)K, N, N, 1.0, AX.data(), (int)K, Cx, ldc, 1.0, T.data(), (int)K);
            std::copy(T.begin(), T.end(), AX.begin());
            maxres = 0.0;
            for (int j = 0; j < N; ++j) {
                double s = 0.0;
                for (int64_t i = 0; i < K; ++i) {
                    const double r = xs(AX.data(), j)[i] - eig[j] * xs(X, j)[i];
                    xs(R.data(), j)[i] = r;
                    s += r * r;
                }
                maxres = std::max(maxres, std::sqrt(s));
            }
            // NextMatrix_d: recurrences for PtP, XtP, XtX (and the H counterparts), then
            // XtR = XtHX - XtX E and RtP = XtHP - E XtP.
            const int nn = N * N;
            for (int pass = 0; pass < 2; ++pass) {
                std::vector<double>& S = pass ? Sa : Sb;
                double* XX = blk(S, 0, 0); double* XR = blk(S, 0, 1); double* XP = blk(S, 0, 2);
                double* RR = blk(S, 1, 1); double* RP = blk(S, 1, 2); double* PP = blk(S, 2, 2);
                double* t1 = S.data() + 2 * nn9;
                double* tPP = t1 + nn; double* tXP = t1 + 2 * nn; double* tXX = t1 + 3 * nn;
                gemm(false, false, N, N, N, 1.0, RR, n3, Cr, ldc, 0.0, t1, N);
                if (use_p) gemm(false, false, N, N, N, 1.0, RP, n3, Cp, ldc, 1.0, t1, N);
                gemm(true, false, N, N, N, 1.0, Cr, ldc, t1, N, 0.0, tPP, N);
                if (use_p) {
                    gemm(true, false, N, N, N, 1.0, RP, n3, Cr, ldc, 0.0, t1, N);
                    gemm(false, false, N, N, N, 1.0, PP, n3, Cp, ldc, 1.0, t1, N);
                    gemm(true, false, N, N, N, 1.0, Cp, ldc, t1, N, 1.0, tPP, N);
                }
                gemm(false, false, N, N, N, 1.0, XR, n3, Cr, ldc, 0.0, t1, N);
                if (use_p) gemm(false, false, N, N, N, 1.0, XP, n3, Cp, ldc, 1.0, t1, N);
                copy_sub(N, N, tXP, N, tPP, N);
                gemm(true, false, N, N, N, 1.0, Cx, ldc, t1, N, 1.0, tXP, N);
                gemm(false, false, N, N, N, 1.0, XX, n3, Cx, ldc, 1.0, t1, N);
                copy_sub(N, N, tXX, N, tXP, N);
                gemm(true, false, N, N, N, 1.0, t1, N, Cx, ldc, 1.0, tXX, N);
                copy_sub(N, N, PP, n3, tPP, N);
                copy_sub(N, N, XP, n3, tXP, N);
                copy_sub(N, N, XX, n3, tXX, N);
            }
            {
                double* XR = blk(Sb, 0, 1);
                const double* tXHX = Sa.data() + 2 * nn9 + 3 * nn;
                const double* tXX = Sb.data() + 2 * nn9 + 3 * nn;
                for (int j = 0; j < N; ++j)
                    for (int i = 0; i < N; ++i) XR[i + (int64_t)n3 * j] = tXHX[i + N * j] - tXX[i + N * j] * eig[j];
                double* RP = blk(Sb, 1, 2);
                const double* tXHP = Sa.data() + 2 * nn9 + 2 * nn;
                const double* tXP = Sb.data() + 2 * nn9 + 2 * nn;
                for (int j = 0; j < N; ++j)
                    for (int i = 0; i < N; ++i) RP[i + (int64_t)n3 * j] = tXHP[i + N * j] - eig[i] * tXP[i + N * j];
            }
            if (tol > 0.0 && maxres <= tol) stop = true;
        }
        if (stop) {
            for (int j = 0; j < N; ++j) {
                const double c = 1.0 / std::sqrt(nx[j] * dvol);
                for (int64_t i = 0; i < K; ++i) xs(X, j)[i] *= c;
            }
            ++iter;
            break;
        }
    }
    if (out) { out->iterations = iter; out->rollbacks = rollbacks; out->hpsi_calls = hcalls; out->max_residual = maxres; }
    return 0;
}

// Convenience: LOBPCG on the oracle Hamiltonian.
struct OrHamCtx { int n[3]; double h[3]; const double* V; const or_nl* nl; };
static void ham_cb(void* ctx, int nb, const double* x, double* ax) {
    OrHamCtx* c = (OrHamCtx*)ctx;
    or_hamiltonian_block(c->n[0], c->n[1], c->n[2], c->h[0], c->h[1], c->h[2], c->V, c->nl, nb, x, ax);
}
OR_API int or_lobpcg_hamiltonian(int nx, int ny, int nz, double dx, double dy, double dz, const double* V,
                          const or_nl* nl, int N, int iter_max, double tol, double* eig, double* X,
                          double* P, int scf_step, int* iterations, int* rollbacks, double* max_residual) {
    OrHamCtx c{{nx, ny, nz}, {dx, dy, dz}, V, nl};
    OrLobpcgOut o{};
    const int64_t ng = (int64_t)nx * ny * nz;
    int rc = or_lobpcg_d_multi(ng, dx * dy * dz, N, iter_max, tol, eig, X, P, scf_step, ham_cb, &c, &o);
    if (iterations) *iterations = o.iterations;
    if (rollbacks) *rollbacks = o.rollbacks;
    if (max_residual) *max_residual = o.max_residual;
    return rc;
}

// LOBPCG on an explicit dense symmetric matrix A (ng x ng, column-major), dvol = 1.
struct OrDenseCtx { int64_t n; const double* A; };
static void dense_cb(void* ctx, int nb, const double* x, double* ax) {
    OrDenseCtx* c = (OrDenseCtx*)ctx;
    gemm(false, false, (int)c->n, nb, (int)c->n, 1.0, c->A, (int)c->n, x, (int)c->n, 0.0, ax, (int)c->n);
}
OR_API int or_lobpcg_dense(int64_t n, const double* A, int N, int iter_max, double tol, double* eig,
                    double* X, double* P, int scf_step, int* iterations, int* rollbacks, double* max_residual) {
    OrDenseCtx c{n, A};
    OrLobpcgOut o{};
    int rc = or_lobpcg_d_multi(n, 1.0, N, iter_max, tol, eig, X, P, scf_step, dense_cb, &c, &o);
    if (iterations) *iterations = o.iterations;
    if (rollbacks) *rollbacks = o.rollbacks;
    if (max_residual) *max_residual = o.max_residual;
    return rc;
}

// lapack_DSYGVX on a caller matrix pair (upper triangles), for the reference's 8x8 check.
OR_API void or_dsygvx(int n, double* A, double* B, double* w, double* Z) { upper_dsygvx(n, A, B, w, Z); }
