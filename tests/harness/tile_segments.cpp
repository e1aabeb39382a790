// Host harness for tests/test_tile_segments.py: builds the projector geometry of a diamond
// cell with the library's host code (qg_setup.cpp), then both per-(tile, plane) forms the
// fused kernels read - the item tables (point, atom slot, projector values) and the row
// segments - and evaluates the segment form with the kernel's formula in FP32 (displacement
// from the integer part of the image centre in grid units).  Prints the largest difference
// to the tables and the number of (tile, plane, point, slot, slot-coefficient) entries that
// one form has and the other lacks.
#include "../../qumasun_amd/csrc/qg_internal.hpp"
#include <cstdio>
#include <cmath>
#include <map>
#include <tuple>
using namespace qg;
extern "C" int qg_synth_diamond_atoms(int ncell, double a, double* xyz);
int main(int argc, char** argv) {
    const int n = argc > 1 ? atoi(argv[1]) : 40;
    const int ty = argc > 2 ? atoi(argv[2]) : 16;
    const int tz = argc > 3 ? atoi(argv[3]) : 2;
    const double A = 6.74;
    qg_grid g; g.n[0] = g.n[1] = g.n[2] = n; g.h[0] = g.h[1] = g.h[2] = A / n;
    double xyz[24]; qg_synth_diamond_atoms(1, A, xyz);
    int species[8] = {0}; int nproj[1] = {2};
    const double sig = 0.6, amp = std::sqrt(4.0 / (std::sqrt(M_PI) * sig * sig * sig));
    qg_proj_spec sp[2] = {};
    sp[0].l = 0; sp[0].energy = 1.0; sp[0].cutoff = 2.4; sp[0].radial_kind = QG_RADIAL_GAUSSIAN; sp[0].sigma = sig; sp[0].amplitude = amp;
    sp[1] = sp[0]; sp[1].l = 1; sp[1].energy = -0.5;
    NlGeometry geo;
    build_nl_geometry(g, 8, xyz, species, 1, nproj, sp, geo);
    TileNL ti, ts;
    build_tile_nl(g, geo, ty, tz, false, ti);
    build_tile_nl(g, geo, ty, tz, true, ts);
    const int nx = n, nz = n, njs = geo.njs;
    // items: key (tile, z, point, slot, j) -> beta
    std::map<std::tuple<int, int, int, int, int>, double> ref, got;
    for (int t = 0; t < ti.ntile; ++t)
        for (int z = 0; z < nz; ++z) {
            const size_t tp = (size_t)t * nz + z;
            for (int64_t it = ti.item_off[tp]; it < ti.item_off[tp + 1]; ++it) {
                const int pt = ti.item_meta[it] & 0xffff, slot = ti.item_meta[it] >> 16;
                for (int j = 0; j < njs; ++j) ref[{t, z, pt, slot, j}] += ti.item_beta[it * njs + j];
            }
        }
    for (int t = 0; t < ts.ntile; ++t) {
        const int y0 = t * ty, rows = std::min(ty, n - y0);
        for (int z = 0; z < nz; ++z) {
            const int slab = z / tz;
            const int wgi = t * ts.nslab + slab;
            const int i0 = ts.wg_img_off[wgi];
            for (int ly = 0; ly < rows; ++ly) {
                const size_t k = ((size_t)t * nz + z) * ty + ly;
                for (int ix = 0; ix < nx; ++ix) {
                    for (int si = ts.seg_off[k]; si < ts.seg_off[k + 1]; ++si) {
                        const uint32_t sv = ts.seg[si];
                        const int x0 = sv & 0x7ff, x1 = (sv >> 11) & 0x7ff, li = sv >> 22;
                        if (ix < x0 || ix >= x1) continue;
                        const DevImage& im = geo.images[ts.wg_img[i0 + li]];
                        const int slot = ts.wg_img_atom[i0 + li];
                        typedef float TT;
                        TT io[3]; int ci[3];
                        for (int d = 0; d < 3; ++d) { const double u = im.R[d] / g.h[d]; const double nn = std::rint(u); io[d] = (TT)(u - nn); ci[d] = im.shift[d] + (int)nn; }
                        const int iy = y0 + ly;
                        const TT xc = (TT)g.h[0] * ((TT)(ix - ci[0]) - io[0]);
                        const TT yc = (TT)g.h[1] * ((TT)(iy - ci[1]) - io[1]);
                        const TT zc = (TT)g.h[2] * ((TT)(z - ci[2]) - io[2]);
                        const TT rr = xc * xc + (yc * yc + zc * zc);
                        TT yv[3] = {0, 0, 0};
                        if (!((TT)1e-14 > rr)) { const TT ir = (TT)0.48860251190291992 / std::sqrt(rr); yv[0] = yc * ir; yv[1] = zc * ir; yv[2] = xc * ir; }
                        for (int pj = im.proj_begin; pj < im.proj_end; ++pj) {
                            const DevProj& d = geo.projs[pj];
                            const TT rad = (TT)d.amp * std::exp(-(rr * (TT)d.inv_2s2));
                            const int pt = ly * nx + ix;
                            if (d.l == 0) got[{t, z, pt, slot, d.coef_rel}] += rad * (TT)0.28209479177387814;
                            else for (int m = 0; m < 3; ++m) got[{t, z, pt, slot, d.coef_rel + m}] += rad * yv[m];
                        }
                    }
                }
            }
        }
    }
    double maxd = 0; int missing = 0, extra = 0;
    for (auto& [key, v] : ref) {
        auto itg = got.find(key);
        if (itg == got.end()) { if (v != 0.0) ++missing; continue; }
        maxd = std::max(maxd, std::fabs(itg->second - v));
    }
    for (auto& [key, v] : got) if (!ref.count(key) && v != 0.0) ++extra;
    std::printf("n %d ty %d tz %d: items %zu segs %zu max|diff| %.3e missing %d extra %d max_img %d max_seg %d\n", n, ty, tz,
                ref.size(), got.size(), maxd, missing, extra, ts.max_img, ts.max_seg);
    return 0;
}
