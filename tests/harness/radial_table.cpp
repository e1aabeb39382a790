// Host harness for tests/test_radial_table.py: builds the projector tables of one atom with a
// tabulated radial part (QG_RADIAL_TABLE) through the library's host code (qg_setup.cpp,
// build_nl_geometry) and reports, per interpolation branch of RadialGrid2::GetValueByXi
// (RadialGrid2.h:81-136), the largest difference between the radial value recovered from the
// tables and a polynomial in xi = log r that the table samples.
//   argv: degree of the polynomial (0..4), l of the channel (0 or 1)
// Branch classes by interval i = floor((xi - xi_min) / xi_delta) of n table points:
//   rmin (r < r_min: table[0]), lin (i = 0 or n-2), cub (i = 1 or n-3), quint (1 < i < n-3),
//   zero (i >= n-1).
#include "../../qumasun_amd/csrc/qg_internal.hpp"
#include <cmath>
#include <cstdio>
using namespace qg;

static double poly(int deg, double x) {
    const double c[5] = {0.7, -0.45, 0.3, -0.12, 0.05};
    double v = 0.0, p = 1.0;
    for (int k = 0; k <= deg; ++k, p *= x) v += c[k] * p;
    return v;
}

int main(int argc, char** argv) {
    const int deg = argc > 1 ? atoi(argv[1]) : 1;
    const int l = argc > 2 ? atoi(argv[2]) : 0;
    const int n = 28;
    const double L = 6.0;
    qg_grid g;
    g.n[0] = g.n[1] = g.n[2] = n;
    g.h[0] = g.h[1] = g.h[2] = L / n;
    // table: xi_min = log 0.3, 9 points of 0.25; last point at r = 2.22 inside the 2.4 cutoff,
    // so the sphere holds points of every branch, r < r_min and beyond the table
    const int nt = 9;
    const double xi_min = std::log(0.3), dxi = 0.25;
    std::vector<double> tab(nt);
    for (int k = 0; k < nt; ++k) tab[k] = poly(deg, xi_min + k * dxi);
    qg_proj_spec sp{};
    sp.l = l;
    sp.energy = 1.0;
    sp.cutoff = 2.4;
    sp.radial_kind = QG_RADIAL_TABLE;
    sp.n_table = nt;
    sp.xi_min = xi_min;
    sp.xi_delta = dxi;
    sp.r_min = 0.3;
    sp.table = tab.data();
    // a generic position (no grid point closer than ~1e-3 bohr), plus one nucleus on a grid
    // point so r = 0 is sampled
    const double xyz[6] = {2.9137, 3.0411, 2.8213, 10 * g.h[0], 12 * g.h[1], 9 * g.h[2]};
    const int species[2] = {0, 0};
    const int nproj[1] = {1};
    NlGeometry geo;
    build_nl_geometry(g, 2, xyz, species, 1, nproj, &sp, geo);
    const double Y00 = 1.0 / std::sqrt(4.0 * M_PI), c1 = std::sqrt(3.0 / (4.0 * M_PI));
    double err[5] = {0, 0, 0, 0, 0};
    int cnt[5] = {0, 0, 0, 0, 0};
    int outside = 0;
    for (int i = 0; i < (int)geo.images.size(); ++i) {
        const DevImage& im = geo.images[i];
        for (int64_t p = geo.img_tbl_off[i]; p < geo.img_tbl_off[i + 1]; ++p) {
            const int64_t gi = geo.gidx[p];
            const int ix = gi % n, iy = (gi / n) % n, iz = gi / ((int64_t)n * n);
            const double x = g.h[0] * (ix - im.shift[0]) - im.R[0];
            const double y = g.h[1] * (iy - im.shift[1]) - im.R[1];
            const double z = g.h[2] * (iz - im.shift[2]) - im.R[2];
            const double rr = x * x + (y * y + z * z);
            if (rr > sp.cutoff * sp.cutoff) ++outside;
            const double* b = &geo.beta[p * geo.njs];
            double rad;
            if (l == 0) {
                rad = b[0] / Y00;
            } else {
                // the three m values are rad * c1 * (y, z, x) / r: recover rad from the largest
                const double r = std::sqrt(rr);
                const double comp[3] = {y, z, x};
                int m = 0;
                for (int k = 1; k < 3; ++k)
                    if (std::fabs(comp[k]) > std::fabs(comp[m])) m = k;
                if (rr < 1e-14) continue;   // Y_1m(0) = 0: nothing to recover
                rad = b[m] / (c1 * comp[m] / r);
            }
            const double xi = 0.5 * std::log(rr);
            int cls;
            double want;
            if (sp.r_min * sp.r_min > rr) {
                cls = 0;
                want = tab[0];
            } else {
                const int k = (int)std::floor((xi - xi_min) / dxi);
                if (k >= nt - 1) { cls = 4; want = 0.0; }
                else if (k == 0 || k == nt - 2) { cls = 1; want = poly(deg, xi); }
                else if (k == 1 || k == nt - 3) { cls = 2; want = poly(deg, xi); }
                else { cls = 3; want = poly(deg, xi); }
            }
            err[cls] = std::max(err[cls], std::fabs(rad - want));
            ++cnt[cls];
        }
    }
    printf("deg %d l %d rmin %d %.3e lin %d %.3e cub %d %.3e quint %d %.3e zero %d %.3e outside %d\n", deg, l,
           cnt[0], err[0], cnt[1], err[1], cnt[2], err[2], cnt[3], err[3], cnt[4], err[4], outside);
    return 0;
}
