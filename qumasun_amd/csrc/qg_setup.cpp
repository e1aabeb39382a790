// qg_setup.cpp — host-side setup of the MI355X engine: error state, grid handles, the
// nonlocal-projector geometry and the seeded synthetic workloads.  Compiled without
// floating-point contraction so sphere membership is decided by exactly the arithmetic of
// the reference (x = dx*(ix - shift) - Rx, rr = x*x + (y*y + z*z), inside iff !(rc^2 < rr)).
#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <thread>
#include "qg_internal.hpp"

namespace qg {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

hipStream_t default_stream() {
    static std::mutex mu;
    static std::vector<hipStream_t> streams;
    int dev = 0;
    QG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if ((int)streams.size() <= dev) streams.resize(dev + 1, nullptr);
    if (!streams[dev]) QG_HIP(hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking));
    return streams[dev];
}

// GridForSphere (GridSubgrid.h:483-492): the box of grid indices around R that can lie
// within rc, half-open [b, e) per axis.
static void sphere_box(const qg_grid& g, const double R[3], double rc, int b[3], int e[3]) {
    for (int d = 0; d < 3; ++d) {
        b[d] = (int)std::ceil((R[d] - rc) / g.h[d] - 0.5);
        e[d] = (int)std::floor((R[d] + rc) / g.h[d] + 0.5) + 1;
    }
}

// Smallest n with n*y > x (integer division toward zero, as GetOverlapPeriodic's helper).
static int first_multiple_above(int x, int y) { return (x >= 0) ? (x / y) + 1 : ((x + 1) / y); }

// RadialGrid2::GetValueByXi (RadialGrid2.h:81-136): linear in the outer intervals, cubic
// Hermite next to them, quintic Hermite with 4th-order derivative estimates inside.
static double radial_by_xi_host(double xi, const double* v, int n, double xi_min, double xi_delta) {
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

// p(r) of a channel at squared distance rr inside its cutoff
// (PseudoPotOperator_mpi_v2.h:154-166 for tables).
static double radial_host(const qg_proj_spec& s, double rr) {
    if (s.radial_kind == QG_RADIAL_GAUSSIAN) return s.amplitude * std::exp(-rr / (2.0 * s.sigma * s.sigma));
    if (s.r_min * s.r_min > rr) return s.table[0];
    return radial_by_xi_host(0.5 * std::log(rr), s.table, s.n_table, s.xi_min, s.xi_delta);
}

// beta_j = p(r) Y_lm on every sphere point, image by image, rows in (z, y) order; real
// Y_lm of RealSphericalHarmonics.h:149-161 (m = -1: y/r, 0: z/r, +1: x/r), Y_lm(r = 0) = 0
// for l > 0 (SetBlockYlm.h:41-42).
static void build_tables(const qg_grid& g, const qg_proj_spec* specs, NlGeometry& out) {
    out.njs = out.max_coef_per_atom <= 4 ? 4 : 8;
    const int nimg = (int)out.images.size();
    out.img_tbl_off.assign(nimg + 1, 0);
    out.row_tbl_off.assign(out.intervals.size(), 0);
    int64_t np = 0;
    for (int i = 0; i < nimg; ++i) {
        const DevImage& im = out.images[i];
        out.img_tbl_off[i] = np;
        const int nrows = (im.e[1] - im.b[1]) * (im.e[2] - im.b[2]);
        for (int r = 0; r < nrows; ++r) {
            const int2 iv = out.intervals[im.row_base + r];
            out.row_tbl_off[im.row_base + r] = np;
            np += std::max(0, iv.y - iv.x);
        }
    }
    out.img_tbl_off[nimg] = np;
    QG_REQUIRE(np < (1LL << 31), "projector tables too large");
    out.beta.assign((size_t)np * out.njs, 0.0);
    out.gidx.assign((size_t)np, 0);
    const double Y00 = 1.0 / std::sqrt(4.0 * M_PI);
    const double c1 = std::sqrt(3.0 * 1.0 / (2.0 * 2.0 * M_PI));   // |m| = 1
    const double c0 = std::sqrt(3.0 / (4.0 * M_PI));               // m = 0
    auto work = [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i) {
            const DevImage& im = out.images[i];
            const int wy = im.e[1] - im.b[1];
            for (int iz = im.b[2]; iz < im.e[2]; ++iz) {
                const double z = g.h[2] * (double)(iz - im.shift[2]) - im.R[2];
                const double zz = z * z;
                for (int iy = im.b[1]; iy < im.e[1]; ++iy) {
                    const double y = g.h[1] * (double)(iy - im.shift[1]) - im.R[1];
                    const double yy_zz = y * y + zz;
                    const int row = im.row_base + (iy - im.b[1]) + wy * (iz - im.b[2]);
                    const int2 iv = out.intervals[row];
                    int64_t pt = out.row_tbl_off[row];
                    for (int ix = iv.x; ix < iv.y; ++ix, ++pt) {
                        const double x = g.h[0] * (double)(ix - im.shift[0]) - im.R[0];
                        const double rr = x * x + yy_zz;
                        out.gidx[pt] = ix + g.n[0] * (iy + g.n[1] * iz);
                        double* b = &out.beta[(size_t)pt * out.njs];
                        for (int pj = im.proj_begin; pj < im.proj_end; ++pj) {
                            const DevProj& d = out.projs[pj];
                            const double rad = radial_host(specs[d.spec_index], rr);
                            if (d.l == 0) {
                                b[d.coef_rel] = rad * Y00;
                            } else {
                                double ym = 0.0, y0 = 0.0, yp = 0.0;
                                if (!(1.0e-14 > rr)) {
                                    const double r = std::sqrt(rr);
                                    ym = c1 * 1.0 * (y / r);
                                    y0 = c0 * (z / r) * 1.0;
                                    yp = c1 * 1.0 * (x / r);
                                }
                                b[d.coef_rel] = rad * ym;
                                b[d.coef_rel + 1] = rad * y0;
                                b[d.coef_rel + 2] = rad * yp;
                            }
                        }
                    }
                }
            }
        }
    };
    const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(work, (int)((int64_t)nimg * t / nt), (int)((int64_t)nimg * (t + 1) / nt));
    for (auto& t : th) t.join();
}

void build_nl_geometry(const qg_grid& g, int natoms, const double* xyz, const int* species,
                       int nspecies, const int* nproj, const qg_proj_spec* specs, NlGeometry& out) {
    out = NlGeometry{};
    out.natoms = natoms;
    std::vector<int> first(nspecies + 1, 0);
    for (int s = 0; s < nspecies; ++s) {
        QG_REQUIRE(nproj[s] >= 0, "negative projector count");
        first[s + 1] = first[s] + nproj[s];
    }
    // Per species: groups of projectors sharing one cutoff (the reference builds one block
    // set per l from cutoff_r[l]; l values with equal cutoffs share identical blocks).
    struct Group { double rc; int pbeg, pend; };
    std::vector<std::vector<Group>> groups(nspecies);
    std::vector<int> coef_per_species(nspecies, 0);
    for (int s = 0; s < nspecies; ++s) {
        std::vector<int> coef_rel(nproj[s]);
        int c = 0;
        for (int k = 0; k < nproj[s]; ++k) {
            const qg_proj_spec& p = specs[first[s] + k];
            QG_REQUIRE(p.l == 0 || p.l == 1, "only l = 0 and l = 1 projectors are supported");
            QG_REQUIRE(p.cutoff > 0.0, "projector cutoff must be positive");
            QG_REQUIRE(p.radial_kind == QG_RADIAL_GAUSSIAN || p.radial_kind == QG_RADIAL_TABLE,
                       "unknown radial kind");
            if (p.radial_kind == QG_RADIAL_TABLE)
                QG_REQUIRE(p.table != nullptr && p.n_table >= 5 && p.xi_delta > 0.0,
                           "tabulated projector needs >= 5 values and xi_delta > 0");
            else
                QG_REQUIRE(p.sigma > 0.0, "gaussian projector needs sigma > 0");
            coef_rel[k] = c;
            c += 2 * p.l + 1;
        }
        coef_per_species[s] = c;
        out.max_coef_per_atom = std::max(out.max_coef_per_atom, c);
        std::vector<double> cutoffs;
        for (int k = 0; k < nproj[s]; ++k) {
            const double rc = specs[first[s] + k].cutoff;
            if (std::find(cutoffs.begin(), cutoffs.end(), rc) == cutoffs.end()) cutoffs.push_back(rc);
        }
        for (double rc : cutoffs) {
            Group gr{rc, (int)out.projs.size(), 0};
            for (int k = 0; k < nproj[s]; ++k) {
                const qg_proj_spec& p = specs[first[s] + k];
                if (p.cutoff != rc) continue;
                DevProj d{};
                d.l = p.l;
                d.kind = p.radial_kind;
                d.coef_rel = coef_rel[k];
                d.spec_index = first[s] + k;
                d.energy = p.energy;
                d.amp = p.amplitude;
                d.inv_2s2 = (p.radial_kind == QG_RADIAL_GAUSSIAN) ? 1.0 / (2.0 * p.sigma * p.sigma) : 0.0;
                d.xi_min = p.xi_min;
                d.xi_delta = p.xi_delta;
                d.r_min2 = p.r_min * p.r_min;
                d.table_off = (int)out.tables.size();
                d.n_table = (p.radial_kind == QG_RADIAL_TABLE) ? p.n_table : 0;
                if (p.radial_kind == QG_RADIAL_TABLE) out.tables.insert(out.tables.end(), p.table, p.table + p.n_table);
                out.projs.push_back(d);
            }
            gr.pend = (int)out.projs.size();
            groups[s].push_back(gr);
        }
    }

    out.atom_coef_off.assign(natoms + 1, 0);
    out.atom_img_off.assign(natoms + 1, 0);
    for (int a = 0; a < natoms; ++a) {
        const int s = species[a];
        QG_REQUIRE(s >= 0 && s < nspecies, "species index out of range");
        out.atom_coef_off[a + 1] = out.atom_coef_off[a] + coef_per_species[s];
        const double R[3] = {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
        for (const Group& gr : groups[s]) {
            int bb[3], be[3];
            sphere_box(g, R, gr.rc, bb, be);
            int ns[3], ne[3];
            for (int d = 0; d < 3; ++d) {
                ns[d] = first_multiple_above(0 - be[d], g.n[d]);
                ne[d] = -first_multiple_above(bb[d] - g.n[d], g.n[d]);
            }
            const double rc2 = gr.rc * gr.rc;
            for (int kz = ns[2]; kz <= ne[2]; ++kz)
                for (int ky = ns[1]; ky <= ne[1]; ++ky)
                    for (int kx = ns[0]; kx <= ne[0]; ++kx) {
                        const int k[3] = {kx, ky, kz};
                        DevImage im{};
                        for (int d = 0; d < 3; ++d) {
                            im.shift[d] = k[d] * g.n[d];
                            im.b[d] = std::max(0, bb[d] + im.shift[d]);
                            im.e[d] = std::min(g.n[d], be[d] + im.shift[d]);
                            im.R[d] = R[d];
                        }
                        if (im.b[0] >= im.e[0] || im.b[1] >= im.e[1] || im.b[2] >= im.e[2]) continue;
                        im.row_base = (int)out.intervals.size();
                        im.coef_base = out.atom_coef_off[a];
                        im.proj_begin = gr.pbeg;
                        im.proj_end = gr.pend;
                        im.atom = a;
                        // exact membership per row, in the reference's arithmetic
                        for (int iz = im.b[2]; iz < im.e[2]; ++iz) {
                            const double z = g.h[2] * (double)(iz - im.shift[2]) - R[2];
                            const double zz = z * z;
                            for (int iy = im.b[1]; iy < im.e[1]; ++iy) {
                                const double y = g.h[1] * (double)(iy - im.shift[1]) - R[1];
                                const double yy_zz = y * y + zz;
                                int x0 = -1, x1 = -1;
                                for (int ix = im.b[0]; ix < im.e[0]; ++ix) {
                                    const double x = g.h[0] * (double)(ix - im.shift[0]) - R[0];
                                    const double rr = x * x + yy_zz;
                                    const bool inside = !(rc2 < rr);
                                    if (inside) {
                                        if (x0 < 0) x0 = ix;
                                        QG_REQUIRE(x1 < 0 || x1 == ix, "non-contiguous sphere row");
                                        x1 = ix + 1;
                                    }
                                }
                                if (x0 < 0) { x0 = 0; x1 = 0; }
                                out.intervals.push_back(make_int2(x0, x1));
                            }
                        }
                        out.images.push_back(im);
                    }
        }
        out.atom_img_off[a + 1] = (int)out.images.size();
    }
    out.ncoef = out.atom_coef_off[natoms];
    out.coef_energy.assign(out.ncoef, 0.0);
    for (int a = 0; a < natoms; ++a) {
        const int s = species[a];
        int c = out.atom_coef_off[a];
        for (int k = 0; k < nproj[s]; ++k) {
            const qg_proj_spec& p = specs[first[s] + k];
            for (int m = 0; m < 2 * p.l + 1; ++m) out.coef_energy[c++] = p.energy;
        }
    }
    build_tables(g, specs, out);
}

void build_tile_nl(const qg_grid& g, const NlGeometry& geo, int ty, int tz, bool segments, TileNL& out) {
    out = TileNL{};
    const int ny = g.n[1], nz = g.n[2];
    out.ty = ty;
    out.tz = tz;
    out.ntile = (ny + ty - 1) / ty;
    out.nslab = (nz + tz - 1) / tz;
    const int nwg = out.ntile * out.nslab;
    // images / atoms per (tile, slab)
    std::vector<std::vector<int>> wimg(nwg), watom(nwg);
    for (int i = 0; i < (int)geo.images.size(); ++i) {
        const DevImage& im = geo.images[i];
        const int t0 = im.b[1] / ty, t1 = (im.e[1] - 1) / ty;
        const int s0 = im.b[2] / tz, s1 = (im.e[2] - 1) / tz;
        for (int t = t0; t <= t1; ++t)
            for (int sl = s0; sl <= s1; ++sl) wimg[t * out.nslab + sl].push_back(i);
    }
    out.wg_img_off.assign(nwg + 1, 0);
    out.wg_atom_off.assign(nwg + 1, 0);
    std::vector<std::vector<int>> local_slot(nwg);
    for (int w = 0; w < nwg; ++w) {
        std::vector<int>& at = watom[w];
        for (int i : wimg[w]) {
            const int a = geo.images[i].atom;
            auto it = std::find(at.begin(), at.end(), a);
            int slot = (int)(it - at.begin());
            if (it == at.end()) at.push_back(a);
            out.wg_img.push_back(i);
            out.wg_img_atom.push_back(slot);
        }
        out.wg_img_off[w + 1] = (int)out.wg_img.size();
        out.wg_atom.insert(out.wg_atom.end(), at.begin(), at.end());
        out.wg_atom_off[w + 1] = (int)out.wg_atom.size();
        out.max_img = std::max(out.max_img, (int)wimg[w].size());
        out.max_atom = std::max(out.max_atom, (int)at.size());
    }
    QG_REQUIRE(out.max_atom < 65536, "too many atoms per workgroup");
    // work items per (tile, plane): every sphere point of every image reaching the plane
    // (the table form of the projector term; none when the kernel evaluates the projectors)
    const int njs = geo.njs;
    out.item_off.assign((size_t)out.ntile * nz + 1, 0);
    for (int t = 0; t < out.ntile && !segments; ++t) {
        const int y0 = t * ty, rows = std::min(ty, ny - y0);
        QG_REQUIRE((int64_t)ty * g.n[0] <= 65535, "tile too large for 16-bit point indices");
        for (int z = 0; z < nz; ++z) {
            const int w = t * out.nslab + z / tz;
            const size_t tp = (size_t)t * nz + z;
            int64_t items = 0;
            for (int li = 0; li < (int)wimg[w].size(); ++li) {
                const DevImage& im = geo.images[wimg[w][li]];
                if (z < im.b[2] || z >= im.e[2]) continue;
                const int wy = im.e[1] - im.b[1];
                const uint32_t slot = (uint32_t)out.wg_img_atom[out.wg_img_off[w] + li];
                for (int y = std::max(y0, im.b[1]); y < std::min(y0 + rows, im.e[1]); ++y) {
                    const int row = im.row_base + (y - im.b[1]) + wy * (z - im.b[2]);
                    const int2 iv = geo.intervals[row];
                    for (int ix = iv.x; ix < iv.y; ++ix) {
                        const int64_t ti = geo.row_tbl_off[row] + (ix - iv.x);
                        out.item_meta.push_back((uint32_t)((y - y0) * g.n[0] + ix) | (slot << 16));
                        out.item_beta.insert(out.item_beta.end(), geo.beta.begin() + ti * njs,
                                             geo.beta.begin() + (ti + 1) * njs);
                        ++items;
                    }
                }
            }
            out.item_off[tp + 1] = out.item_off[tp] + items;
            out.max_items = std::max<int>(out.max_items, (int)items);
        }
    }
    // sphere segments per (tile, plane, row), for the in-kernel projector evaluation
    // (11-bit x bounds, 10-bit image slot: the caller falls back to the tables beyond that)
    if (!segments || g.n[0] > 1024 || out.max_img >= 1024) return;
    out.seg_off.assign((size_t)out.ntile * nz * ty + 1, 0);
    for (int t = 0; t < out.ntile; ++t) {
        const int y0 = t * ty, rows = std::min(ty, ny - y0);
        for (int z = 0; z < nz; ++z) {
            const int w = t * out.nslab + z / tz;
            for (int r = 0; r < ty; ++r) {
                const size_t k = ((size_t)t * nz + z) * ty + r;
                int cnt = 0;
                if (r < rows) {
                    const int y = y0 + r;
                    for (int li = 0; li < (int)wimg[w].size(); ++li) {
                        const DevImage& im = geo.images[wimg[w][li]];
                        if (z < im.b[2] || z >= im.e[2] || y < im.b[1] || y >= im.e[1]) continue;
                        const int row = im.row_base + (y - im.b[1]) + (im.e[1] - im.b[1]) * (z - im.b[2]);
                        const int2 iv = geo.intervals[row];
                        if (iv.y <= iv.x) continue;
                        out.seg.push_back((uint32_t)iv.x | ((uint32_t)iv.y << 11) | ((uint32_t)li << 22));
                        ++cnt;
                    }
                }
                out.seg_off[k + 1] = out.seg_off[k] + cnt;
            }
        }
        for (int sl = 0; sl < out.nslab; ++sl) {
            const int zb = sl * tz, ze = std::min(zb + tz, nz);
            const int n = out.seg_off[((size_t)t * nz + ze) * ty] - out.seg_off[((size_t)t * nz + zb) * ty];
            out.max_seg = std::max(out.max_seg, n);
        }
    }
}

// Build the tile-column images, the (tile, plane) items and the projection's thread map for
// tiles of ty rows and projection slabs of tzp planes.  Returns false when the geometry does
// not fit the encodings (then the caller uses the table kernels).
bool build_rows_geometry(const qg_grid& g, const NlGeometry& geo, int ty, int tzp, bool by_group, RowsGeo& o) {
    // o is freshly constructed by the caller
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    if (nx > 1024 || ty > 32) return false;
    o.ty = ty;
    o.tzp = tzp;
    o.by_group = by_group;
    o.ng4 = (nx + 3) / 4;
    o.ntile = (ny + ty - 1) / ty;
    o.nslabp = (nz + tzp - 1) / tzp;
    // images of each tile column
    o.col_off.assign(o.ntile + 1, 0);
    std::vector<std::vector<int>> col(o.ntile);
    for (int i = 0; i < (int)geo.images.size(); ++i) {
        const DevImage& im = geo.images[i];
        for (int t = im.b[1] / ty; t <= (im.e[1] - 1) / ty; ++t) col[t].push_back(i);
    }
    for (int t = 0; t < o.ntile; ++t) {
        if (col[t].size() >= 1024) return false;
        o.col_img.insert(o.col_img.end(), col[t].begin(), col[t].end());
        o.col_off[t + 1] = (int)o.col_img.size();
        o.max_col = std::max(o.max_col, (int)col[t].size());
    }
    if (by_group) {
        // items cut at multiples of 4 in x, sorted by (row, group, image); per (tile, plane)
        // the run offsets of every (row, group)
        o.it_off.assign((size_t)o.ntile * nz + 1, 0);
        const int ng4 = o.ng4;
        o.grp_off.assign((size_t)o.ntile * nz * (ty * ng4 + 1), 0);
        std::vector<std::pair<uint64_t, uint32_t>> pl;   // (sort key, item) of one (tile, plane)
        for (int t = 0; t < o.ntile; ++t) {
            const int y0 = t * ty, rows = std::min(ty, ny - y0);
            const int nc = (int)col[t].size();
            for (int z = 0; z < nz; ++z) {
                pl.clear();
                for (int li = 0; li < nc; ++li) {
                    const DevImage& im = geo.images[col[t][li]];
                    if (z < im.b[2] || z >= im.e[2]) continue;
                    const int wy = im.e[1] - im.b[1];
                    for (int y = std::max(y0, im.b[1]); y < std::min(y0 + rows, im.e[1]); ++y) {
                        const int2 iv = geo.intervals[im.row_base + (y - im.b[1]) + wy * (z - im.b[2])];
                        for (int xs = iv.x; xs < iv.y;) {
                            const int xe = std::min(iv.y, (xs / 4 + 1) * 4);
                            const uint32_t it = (uint32_t)(y - y0) | ((uint32_t)xs << 5) |
                                                ((uint32_t)(xe - xs - 1) << 15) | ((uint32_t)li << 17);
                            const uint64_t key = ((uint64_t)((y - y0) * ng4 + xs / 4) << 32) | (uint32_t)li;
                            pl.push_back({key, it});
                            xs = xe;
                        }
                    }
                }
                std::sort(pl.begin(), pl.end());
                const size_t tp = (size_t)t * nz + z;
                int* go = &o.grp_off[tp * (ty * ng4 + 1)];
                size_t k = 0;
                for (int gi = 0; gi < ty * ng4; ++gi) {
                    go[gi] = (int)k;
                    while (k < pl.size() && (int)(pl[k].first >> 32) == gi) ++k;
                }
                go[ty * ng4] = (int)k;
                for (auto& e : pl) o.items.push_back(e.second);
                const int n = (int)pl.size();
                o.max_items = std::max(o.max_items, n);
                o.it_off[tp + 1] = o.it_off[tp] + n;
                if ((int64_t)o.it_off[tp] + n >= (1LL << 31) - 1) return false;
            }
        }
        o.atom_slot_off.assign(geo.natoms + 1, 0);
    } else {
        // items per (tile, plane), grouped by column image, with per (plane, image) offsets
        o.it_off.assign((size_t)o.ntile * nz + 1, 0);
        o.pl_base.assign(o.ntile + 1, 0);
        for (int t = 0; t < o.ntile; ++t) o.pl_base[t + 1] = o.pl_base[t] + nz * ((int)col[t].size() + 1);
        o.pl_off.assign((size_t)o.pl_base[o.ntile], 0);
        // per (tile, projection slab, image): number of items
        std::vector<std::vector<int>> slab_cnt((size_t)o.ntile * o.nslabp);
        for (int t = 0; t < o.ntile; ++t) {
            const int y0 = t * ty, rows = std::min(ty, ny - y0);
            const int nc = (int)col[t].size();
            for (int s = 0; s < o.nslabp; ++s) slab_cnt[(size_t)t * o.nslabp + s].assign(nc, 0);
            for (int z = 0; z < nz; ++z) {
                const size_t tp = (size_t)t * nz + z;
                int* po = &o.pl_off[o.pl_base[t] + (size_t)z * (nc + 1)];
                int n = 0;
                for (int li = 0; li < nc; ++li) {
                    po[li] = n;
                    const DevImage& im = geo.images[col[t][li]];
                    if (z < im.b[2] || z >= im.e[2]) continue;
                    const int wy = im.e[1] - im.b[1];
                    for (int y = std::max(y0, im.b[1]); y < std::min(y0 + rows, im.e[1]); ++y) {
                        const int2 iv = geo.intervals[im.row_base + (y - im.b[1]) + wy * (z - im.b[2])];
                        for (int xs = iv.x; xs < iv.y; xs += 4) {
                            const int c = std::min(4, iv.y - xs);
                            o.items.push_back((uint32_t)(y - y0) | ((uint32_t)xs << 5) | ((uint32_t)(c - 1) << 15) |
                                              ((uint32_t)li << 17));
                            ++n;
                            ++slab_cnt[(size_t)t * o.nslabp + z / tzp][li];
                        }
                    }
                }
                po[nc] = n;
                o.max_items = std::max(o.max_items, n);
                o.it_off[tp + 1] = o.it_off[tp] + n;
                if ((int64_t)o.it_off[tp] + n >= (1LL << 31) - 1) return false;
            }
        }
        // projection thread map: every active image of a (tile, slab) owns a contiguous range
        // of threads in proportion to its items; each range gets one partial slot
        constexpr int T = kRowsProjThreads;
        o.thr_map.assign((size_t)o.ntile * o.nslabp * T, 0xffffffffu);
        o.thr_slot.assign((size_t)o.ntile * o.nslabp * T, 0);
        std::vector<std::vector<int>> atom_slots(geo.natoms);
        int nslot = 0;
        for (int t = 0; t < o.ntile; ++t)
            for (int s = 0; s < o.nslabp; ++s) {
                const std::vector<int>& cnt = slab_cnt[(size_t)t * o.nslabp + s];
                std::vector<int> act;
                int64_t tot = 0;
                for (int li = 0; li < (int)cnt.size(); ++li)
                    if (cnt[li] > 0) {
                        act.push_back(li);
                        tot += cnt[li];
                    }
                if (act.empty()) continue;
                if ((int)act.size() > T) return false;
                std::vector<int> nthr(act.size());
                int sum = 0;
                for (size_t k = 0; k < act.size(); ++k) {
                    nthr[k] = std::max(1, (int)((int64_t)T * cnt[act[k]] / tot));
                    sum += nthr[k];
                }
                while (sum > T) {   // take from the image with the most threads
                    size_t m = 0;
                    for (size_t k = 1; k < act.size(); ++k)
                        if (nthr[k] > nthr[m]) m = k;
                    --nthr[m];
                    --sum;
                }
                while (sum < T) {   // give to the image with the most items per thread
                    size_t m = 0;
                    for (size_t k = 1; k < act.size(); ++k)
                        if ((int64_t)cnt[act[k]] * nthr[m] > (int64_t)cnt[act[m]] * nthr[k]) m = k;
                    ++nthr[m];
                    ++sum;
                }
                int tb = 0;
                for (size_t k = 0; k < act.size(); ++k) {
                    const int n = std::min(nthr[k], 256);
                    const int slot = nslot++;
                    atom_slots[geo.images[col[t][act[k]]].atom].push_back(slot);
                    for (int u = 0; u < n; ++u) {
                        const size_t at = ((size_t)t * o.nslabp + s) * T + tb + u;
                        o.thr_map[at] = (uint32_t)act[k] | ((uint32_t)u << 10) | ((uint32_t)(n - 1) << 18);
                        o.thr_slot[at] = slot;
                    }
                    tb += nthr[k];
                }
            }
        o.nslot = nslot;
        o.atom_slot_off.assign(geo.natoms + 1, 0);
        for (int a = 0; a < geo.natoms; ++a) {
            o.atom_slot.insert(o.atom_slot.end(), atom_slots[a].begin(), atom_slots[a].end());
            o.atom_slot_off[a + 1] = (int)o.atom_slot.size();
        }
    }   // projection order
    // s + p layout everywhere: every image holds an s channel (slot 0) and a p channel
    // (slots 1..3), both Gaussian
    o.sp4 = geo.max_coef_per_atom == 4;
    for (const DevImage& im : geo.images) {
        if (im.proj_end - im.proj_begin != 2) { o.sp4 = false; break; }
        const DevProj& s = geo.projs[im.proj_begin];
        const DevProj& p = geo.projs[im.proj_begin + 1];
        if (!(s.l == 0 && s.coef_rel == 0 && p.l == 1 && p.coef_rel == 1)) { o.sp4 = false; break; }
    }
    o.njs = geo.max_coef_per_atom <= 4 ? 4 : 8;
    // s + p with one Gaussian exponent for both channels of every image
    bool same = o.sp4;
    for (const DevImage& im : geo.images) {
        if (!same) break;
        same = geo.projs[im.proj_begin].inv_2s2 == geo.projs[im.proj_begin + 1].inv_2s2;
    }
    o.pform = o.sp4 ? (same ? 2 : 1) : 0;
    return true;
}

}  // namespace qg

using namespace qg;

extern "C" {

const char* qg_last_error(void) { return g_last_error.c_str(); }
int qg_version(void) { return 100; }

int qg_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int qg_grid_create(int nx, int ny, int nz, double dx, double dy, double dz, qg_grid** out) {
    return guarded([&] {
        QG_REQUIRE(out, "null output handle");
        QG_REQUIRE(nx >= 4 && ny >= 4 && nz >= 4, "grid must be at least 4 points per axis (4-point halo)");
        QG_REQUIRE(nx <= 1024 && ny <= 4096 && nz <= 4096, "grid axis too large");
        QG_REQUIRE(dx > 0 && dy > 0 && dz > 0, "grid spacings must be positive");
        qg_grid* g = new qg_grid;
        g->n[0] = nx; g->n[1] = ny; g->n[2] = nz;
        g->h[0] = dx; g->h[1] = dy; g->h[2] = dz;
        *out = g;
    });
}

int qg_grid_destroy(qg_grid* g) {
    delete g;
    return QG_OK;
}

int qg_synth_diamond_atoms(int ncell, double a, double* xyz) {
    return guarded([&] {
        QG_REQUIRE(ncell >= 1 && xyz, "bad arguments");
        // the 8 sites of the cubic diamond cell (fractional), input/diamond8.txt:28-35
        static const double site[8][3] = {{0.00, 0.00, 0.00}, {0.25, 0.25, 0.25}, {0.50, 0.50, 0.00},
                                          {0.75, 0.75, 0.25}, {0.50, 0.00, 0.50}, {0.75, 0.25, 0.75},
                                          {0.00, 0.50, 0.50}, {0.25, 0.75, 0.75}};
        int idx = 0;
        for (int cz = 0; cz < ncell; ++cz)
            for (int cy = 0; cy < ncell; ++cy)
                for (int cx = 0; cx < ncell; ++cx)
                    for (int s = 0; s < 8; ++s) {
                        xyz[3 * idx + 0] = (cx + site[s][0]) * a;
                        xyz[3 * idx + 1] = (cy + site[s][1]) * a;
                        xyz[3 * idx + 2] = (cz + site[s][2]) * a;
                        ++idx;
                    }
    });
}

int qg_synth_vloc(int nx, int ny, int nz, double lx, double ly, double lz, int natoms,
                  const double* xyz, double zion, double rc, double noise, uint32_t seed,
                  double* V) {
    return guarded([&] {
        QG_REQUIRE(nx > 0 && ny > 0 && nz > 0 && V && (natoms == 0 || xyz), "bad arguments");
        const int64_t ng = (int64_t)nx * ny * nz;
        std::mt19937 mt(seed);
        std::uniform_real_distribution<> dist(-noise, noise);
        for (int64_t i = 0; i < ng; ++i) V[i] = dist(mt);
        const double hx = lx / nx, hy = ly / ny, hz = lz / nz;
        const double v0 = -zion * 2.0 / (std::sqrt(M_PI) * rc);   // limit r -> 0
        auto work = [&](int z0, int z1) {
            for (int iz = z0; iz < z1; ++iz)
                for (int iy = 0; iy < ny; ++iy)
                    for (int ix = 0; ix < nx; ++ix) {
                        double acc = 0.0;
                        for (int a = 0; a < natoms; ++a) {
                            double d[3] = {ix * hx - xyz[3 * a], iy * hy - xyz[3 * a + 1], iz * hz - xyz[3 * a + 2]};
                            const double L[3] = {lx, ly, lz};
                            for (int k = 0; k < 3; ++k) d[k] -= L[k] * std::nearbyint(d[k] / L[k]);
                            const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                            acc += (r < 1e-12) ? v0 : -zion * std::erf(r / rc) / r;
                        }
                        V[ix + (int64_t)nx * (iy + (int64_t)ny * iz)] += acc;
                    }
        };
        const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) th.emplace_back(work, (int)((int64_t)nz * t / nt), (int)((int64_t)nz * (t + 1) / nt));
        for (auto& t : th) t.join();
    });
}

int qg_synth_random_states(int64_t ng, int nbands, double dvol, uint32_t seed, double* X) {
    return guarded([&] {
        QG_REQUIRE(ng > 0 && nbands > 0 && X && dvol > 0, "bad arguments");
        std::mt19937 mt(seed);
        std::uniform_real_distribution<> dist(0.0, 1.0);
        for (int n = 0; n < nbands; ++n) {
            double* x = X + (int64_t)n * ng;
            for (int64_t i = 0; i < ng; ++i) x[i] = dist(mt);
            double s = 0.0;
            for (int64_t i = 0; i < ng; ++i) s += x[i] * x[i];
            const double c = 1.0 / std::sqrt(s * dvol);
            for (int64_t i = 0; i < ng; ++i) x[i] *= c;
            (void)dist(mt);   // the eigenvalue-guess draw of the reference's sequence
        }
    });
}

}  // extern "C"
