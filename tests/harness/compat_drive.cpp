// Driver for tests/test_gpu_compat.py: a QUMASUN-style program written against the
// reference-shaped C++ surface (include/qumasun_compat.hpp) only, on host arrays, as a
// QUMASUN call site would be after switching to the engine.  The 8-atom diamond cell on a 24^3
// grid (the seeded workload of the Python package: Vloc = sum -4 erf(r/0.8)/r + U(+-0.05),
// Gaussian s + p projectors), then:
//   1. Laplasian8th(l_grid, out, p, -1/2, dx, dy, dz) on band 0 (out starts at zero)
//   2. ProjectionPP(Hp, psi, l_grid, nuclei, n) on band 0 (Hp starts at zero)
//   3. HamiltonianGPU::Apply on bands 0..3
//   4. LOBPCG::Eigen_d_multi_mpi(l_grid, dVol, 8, 12, ev, X, keep_P, work, nullptr, 0, &H, watch)
//   5. the same 12 iterations with a per-band user operator OpeA(Ax, x) on host pointers
//   6. UpdatePosition with moved nuclei, then Hamiltonian Apply on band 0
//   7. EigenLOBPCG_d_multi(N, dVol, 8, 40, ev, X, keep, work, true, &H, watch)
//   8. KpointHamiltonianGPU(H, k) as OpeA(Ax, x) on one SoAComplex band, then
//      LOBPCG::Eigen_z_multi on 4 planar complex bands
//   9. Gradient8th_ddm, EnergyPPnonlocal, ForceNonlocal on band 0 and EnergyForceNonlocal on
//      the block
// Writes every result to argv[1] (raw float64, in that order) and prints the sector counts
// of the watch.
#include "qumasun_compat.hpp"
#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

using namespace qumasun_gpu;

// The reference's StopWatch interface (StopWatch.h:20-45): Record(idx, count) attributes the
// time since the previous Record to sector idx.
struct Watch {
    double dur[32] = {0};
    int64_t cnt[32] = {0};
    std::chrono::steady_clock::time_point prev = std::chrono::steady_clock::now();
    double Record(int idx, int64_t count = 1) {
        const auto now = std::chrono::steady_clock::now();
        const double d = std::chrono::duration<double>(now - prev).count();
        dur[idx] += d;
        cnt[idx] += count;
        prev = now;
        return d;
    }
};

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: compat_drive OUT [VPS]\n");
        return 2;
    }
    const int n = 24;
    const double a = 6.74, h = a / n, dvol = h * h * h;
    const size_t ng = (size_t)n * n * n;
    const int nb = 8;
    GridRange g{0, 0, 0, n, n, n};
    double xyz[24];
    check(qg_synth_diamond_atoms(1, a, xyz), "atoms");
    std::vector<double> V(ng);
    check(qg_synth_vloc(n, n, n, a, a, a, 8, xyz, 4.0, 0.8, 0.05, 123456790u, V.data()), "vloc");
    std::vector<double> X0(ng * nb);
    check(qg_synth_random_states((int64_t)ng, nb, dvol, 123456789u, X0.data()), "states");
    const double sig = 0.6, amp = std::sqrt(4.0 / (std::sqrt(M_PI) * sig * sig * sig));
    qg_proj_spec s{}, p{};
    s.l = 0; s.energy = 1.0; s.cutoff = 2.4; s.radial_kind = QG_RADIAL_GAUSSIAN; s.sigma = sig; s.amplitude = amp;
    p = s; p.l = 1; p.energy = -0.5;
    Nucleus nuc[8];
    for (int i = 0; i < 8; ++i) nuc[i] = Nucleus{6, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};

    std::FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 3;
    auto put = [&](const double* v, size_t cnt) { std::fwrite(v, sizeof(double), cnt, f); };

    PseudoPotIntegratorGPU pp;
    pp.InitializeGrid(g, h, h, h);
    pp.Load(6, {s, p});
    pp.UpdatePosition(nuc, 8, g);
    HamiltonianGPU H(pp, V.data());

    std::vector<double> out(ng, 0.0);
    Laplasian8th<double>(g, out.data(), X0.data(), -0.5, h, h, h);
    put(out.data(), ng);
    std::fill(out.begin(), out.end(), 0.0);
    pp.ProjectionPP(out.data(), X0.data(), g, nuc, 8);
    put(out.data(), ng);
    std::vector<double> hp(ng * 4);
    H.Apply<double>(4, hp.data(), X0.data());
    put(hp.data(), ng * 4);

    Watch watch;
    std::vector<double> X = X0, P(ng * nb, 0.0), work(LOBPCG::WorkSize_d_multi_mpi((int)ng, nb)), ev(nb);
    LOBPCG::Eigen_d_multi_mpi(g, dvol, nb, 12, ev.data(), X.data(), P.data(), work.data(), nullptr, 0, &H, watch);
    put(ev.data(), nb);
    put(X.data(), ng * nb);

    std::vector<double> X2 = X0, P2(ng * nb, 0.0), ev2(nb);
    Watch w2;
    auto OpeA = [&H](double* Ax, double* x) { H.Apply<double>(1, Ax, x); };
    LOBPCG::Eigen_d_multi_mpi(g, dvol, nb, 12, ev2.data(), X2.data(), P2.data(), work.data(), nullptr, 0, OpeA, w2);
    put(ev2.data(), nb);

    Nucleus moved[8];
    for (int i = 0; i < 8; ++i) moved[i] = Nucleus{6, nuc[i].Rx + 0.11 * (i % 3), nuc[i].Ry - 0.07 * (i % 2), nuc[i].Rz + 0.05};
    pp.UpdatePosition(moved, 8, g);
    H.Apply<double>(1, hp.data(), X0.data());
    put(hp.data(), ng);

    pp.UpdatePosition(nuc, 8, g);
    std::vector<double> X3 = X0, K3(ng * nb, 0.0), ev3(nb);
    Watch w3;
    EigenLOBPCG_d_multi((int)ng, dvol, nb, 40, ev3.data(), X3.data(), K3.data(), work.data(), true, &H, w3);
    put(ev3.data(), nb);

    {
        // 8. the k-point operator through the reference's OpeA(Ax, x) on one SoAComplex band
        //    (re = band 0, im = band 1 of the start), then a 4-band complex solve on planar bands
        //    (re = band b, im = band b + 4)
        struct SoA {
            double* re;
            double* im;
        };
        KpointHamiltonianGPU Hk(H, 0.3, -0.2, 0.25);
        std::vector<double> xr(X0.begin(), X0.begin() + ng), xi(X0.begin() + ng, X0.begin() + 2 * ng), ar(ng), ai(ng);
        SoA x{xr.data(), xi.data()}, ax{ar.data(), ai.data()};
        Hk(ax, x);
        put(ar.data(), ng);
        put(ai.data(), ng);
        const int nk = 4;
        std::vector<double> Z(2 * ng * nk), evk(nk);
        for (int b = 0; b < nk; ++b) {
            std::copy(X0.begin() + b * ng, X0.begin() + (b + 1) * ng, Z.begin() + 2 * b * ng);
            std::copy(X0.begin() + (b + 4) * ng, X0.begin() + (b + 5) * ng, Z.begin() + (2 * b + 1) * ng);
        }
        LOBPCG::Eigen_z_multi(Hk, nk, 500, evk.data(), Z.data(), 1e-8);
        put(evk.data(), nk);
        put(Z.data(), 2 * ng * nk);
    }

    {
        // 9. nonlocal energy and forces: Gradient8th_ddm of band 0 scaled by -2 occ, then
        //    EnergyPPnonlocal and ForceNonlocal on band 0 (occ 0.7), and EnergyForceNonlocal on
        //    the 8-band block with occupancies 2, 2, 1.5, 1, 0.5, 0, 0.25, 0
        const double occ0 = 0.7;
        std::vector<double> grad(3 * ng), fk(24, 0.0);
        Gradient8th_ddm(g, grad.data(), grad.data() + ng, grad.data() + 2 * ng, X0.data(), -2.0 * occ0, h, h, h);
        put(grad.data(), 3 * ng);
        const double e0 = pp.EnergyPPnonlocal(X0.data(), g, nuc, 8);
        const double ef = pp.ForceNonlocal(fk.data(), X0.data(), grad.data(), g, nuc, 8);
        put(&e0, 1);
        put(&ef, 1);
        put(fk.data(), 24);
        const double occ[8] = {2.0, 2.0, 1.5, 1.0, 0.5, 0.0, 0.25, 0.0};
        double eb = 0.0;
        std::vector<double> fb(24);
        pp.EnergyForceNonlocal(8, X0.data(), occ, &eb, fb.data());
        put(&eb, 1);
        put(fb.data(), 24);
    }

    if (argc > 2) {
        // projector channels from an MBK .vps file: Load(Z, filepath, comm), then ProjectionPP
        int comm = 0;
        PseudoPotIntegratorGPU pv;
        pv.InitializeGrid(g, h, h, h);
        pv.Load(6, argv[2], comm);
        pv.UpdatePosition(nuc, 8, g);
        std::fill(out.begin(), out.end(), 0.0);
        pv.ProjectionPP(out.data(), X0.data(), g, nuc, 8);
        put(out.data(), ng);
    }
    std::fclose(f);
    std::printf("watch");
    for (int k = 10; k < 20; ++k) std::printf(" %d:%lld", k, (long long)watch.cnt[k]);
    std::printf("\nwatch_user");
    for (int k = 10; k < 20; ++k) std::printf(" %d:%lld", k, (long long)w2.cnt[k]);
    std::printf("\n");
    return 0;
}
