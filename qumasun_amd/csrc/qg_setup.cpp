// qg_setup.cpp — host-side setup of the MI355X engine: error state, grid handles, the
// nonlocal-projector geometry and the seeded synthetic workloads.  Compiled without
// floating-point contraction so sphere membership is decided by exactly the arithmetic of
// the reference (x = dx*(ix - shift) - Rx, rr = x*x + (y*y + z*z), inside iff !(rc^2 < rr)).
#include <algorithm>
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
}

void build_tile_lists(const qg_grid& g, const NlGeometry& geo, int ty, std::vector<int>& off,
                      std::vector<int>& ids) {
    const int ntile = (g.n[1] + ty - 1) / ty;
    const int nz = g.n[2];
    std::vector<int> cnt((size_t)ntile * nz + 1, 0);
    auto visit = [&](auto&& f) {
        for (int i = 0; i < (int)geo.images.size(); ++i) {
            const DevImage& im = geo.images[i];
            const int t0 = im.b[1] / ty, t1 = (im.e[1] - 1) / ty;
            for (int t = t0; t <= t1; ++t)
                for (int z = im.b[2]; z < im.e[2]; ++z) f((size_t)t * nz + z, i);
        }
    };
    visit([&](size_t slot, int) { ++cnt[slot + 1]; });
    for (size_t i = 1; i < cnt.size(); ++i) cnt[i] += cnt[i - 1];
    off = cnt;
    ids.assign(off.back(), 0);
    std::vector<int> fill(off.begin(), off.end() - 1);
    visit([&](size_t slot, int i) { ids[fill[slot]++] = i; });
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
