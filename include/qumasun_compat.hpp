// qumasun_compat.hpp — the reference's C++ call surface of QUMASUN's H·psi + LOBPCG path on
// top of the C ABI (qumasun_gpu.h), so QUMASUN call sites switch to the MI355X engine with
// the same arguments and meaning.  Header-only; errors become std::runtime_error (the
// reference returns void and prints; the engine reports every failure).
//
//   Laplasian8th<T>(l_grid, out, p, coef, dx, dy, dz)               GridDifference8th.h:177-185
//   PseudoPotIntegratorGPU::InitializeGrid / Load / UpdatePosition   PseudoPotOperator.h:93,
//                                                                    PseudoPotOperator_vlocal.h:80,
//                                                                    PseudoPotOperator_mpi_v2.h:231
//   PseudoPotIntegratorGPU::ProjectionPP(Hp, psi, l_grid, nuclei, n) PseudoPotOperator_mpi_v2.h:465
//   PseudoPotIntegratorGPU::ProjectionPP_bundle                      PseudoPotOperator_mpi_v2.h:789
//   HamiltonianGPU (mHamiltonianMatrix_ddm), usable as OpeA          qumasun_mpi.matrix.h:21-43
//   LOBPCG::Eigen_d_multi_mpi(l_grid, dVol, N, iter_max, ev, X, keep_P, work,
//                             keep_S_matrix, SCF_step, OpeA, watch)  lobpcg_d_multi_mpi.h:882-890
//   EigenLOBPCG_d_multi(N, dVol, num_solution, iter_max, ev, X, keep, work,
//                       is_first_time, OpeA, watch)                  lobpcg_d_multi.h:23-27
//   GramSchmidt(N, num_solution, psi, dVol)                          GramSchmidt.h:6-38
//
// Memory: every array argument may be host or device memory (checked per call with
// hipPointerGetAttributes).  Host arrays are staged through device memory, one copy each way,
// as a drop-in for host-resident code; device arrays are used in place.  A user operator OpeA
// receives pointers in the memory space of the caller's eigen_vectors.
// Grid types: any type with SizeX()/SizeY()/SizeZ()/Size3D() (the reference's GridRange and
// GridRangeMPI both qualify); one process holds the whole periodic grid.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>
#include "qumasun_gpu.h"

namespace qumasun_gpu {

inline void check(int rc, const char* what) {
    if (rc != QG_OK) throw std::runtime_error(std::string(what) + ": " + qg_last_error());
}
inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// The reference's box of grid indices [begin, end) (GridRange.h:14-46).
struct GridRange {
    int begin_x, begin_y, begin_z;
    int end_x, end_y, end_z;
    size_t Size3D() const { return (size_t)SizeX() * SizeY() * SizeZ(); }
    int SizeX() const { return end_x - begin_x; }
    int SizeY() const { return end_y - begin_y; }
    int SizeZ() const { return end_z - begin_z; }
};

// Nucleus{Z, Rx, Ry, Rz} (nucleus.h:3-8)
struct Nucleus {
    int Z;
    double Rx, Ry, Rz;
};

template <class T> struct dtype_of;
template <> struct dtype_of<double> { static constexpr int value = QG_F64; };
template <> struct dtype_of<float> { static constexpr int value = QG_F32; };

// Whether p is device memory (allocations of hipMalloc / hipMallocManaged).
inline bool is_device_ptr(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();   // plain host memory is not registered: clear the error
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Device copy of a caller array, copied back on sync_back() when the caller's is host memory.
template <class T>
class Staged {
public:
    Staged(T* user, size_t n, bool copy_in) : user_(user), n_(n) {
        if (is_device_ptr(user) || n == 0) {
            dev_ = user;
            return;
        }
        hip_check(hipMalloc(&dev_, n * sizeof(T)), "hipMalloc");
        own_ = true;
        if (copy_in) hip_check(hipMemcpy(dev_, user, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    }
    ~Staged() {
        if (own_) (void)hipFree(dev_);
    }
    Staged(const Staged&) = delete;
    Staged& operator=(const Staged&) = delete;
    T* get() const { return dev_; }
    bool host() const { return own_; }
    void sync_back() {
        if (own_) hip_check(hipMemcpy(user_, dev_, n_ * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
    }

private:
    T* user_;
    T* dev_ = nullptr;
    size_t n_;
    bool own_ = false;
};

// RAII grid handle
class Grid {
public:
    Grid(int nx, int ny, int nz, double dx, double dy, double dz) {
        check(qg_grid_create(nx, ny, nz, dx, dy, dz, &h_), "qg_grid_create");
    }
    template <class G>
    Grid(const G& g, double dx, double dy, double dz) : Grid(g.SizeX(), g.SizeY(), g.SizeZ(), dx, dy, dz) {}
    ~Grid() { qg_grid_destroy(h_); }
    Grid(const Grid&) = delete;
    Grid& operator=(const Grid&) = delete;
    qg_grid* get() const { return h_; }

private:
    qg_grid* h_ = nullptr;
};

// Grid handles of the free functions, kept per (shape, spacing) for the life of the process.
inline qg_grid* cached_grid(int nx, int ny, int nz, double dx, double dy, double dz) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, double, double, double>, std::unique_ptr<Grid>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto& g = cache[std::make_tuple(nx, ny, nz, dx, dy, dz)];
    if (!g) g.reset(new Grid(nx, ny, nz, dx, dy, dz));
    return g->get();
}

// out += coef * Lap8 p (periodic) for one band: Laplasian8th<T>(l_grid, out, p, coef, dx, dy,
// dz).  Device arrays run on the null stream without a host wait (HIP order); host arrays are
// staged and the call returns with out updated.
template <class T, class G>
void Laplasian8th(const G& l_grid, T* out, const T* p, double coef, double dx, double dy, double dz) {
    qg_grid* g = cached_grid(l_grid.SizeX(), l_grid.SizeY(), l_grid.SizeZ(), dx, dy, dz);
    const size_t ng = (size_t)l_grid.Size3D();
    Staged<T> so(out, ng, true);
    Staged<T> sp(const_cast<T*>(p), ng, true);
    check(qg_laplacian8_apply(g, dtype_of<T>::value, 1, sp.get(), (int64_t)ng, so.get(), (int64_t)ng, coef, 1,
                              nullptr),
          "qg_laplacian8_apply");
    so.sync_back();
}

// Gradient8th_ddm(l_grid, dpsi_dx, dpsi_dy, dpsi_dz, psi, coef, dx, dy, dz)
// (GridGradient8th.h:17-160): the three coef-scaled 8th-order derivatives of one band (set).
template <class G>
void Gradient8th_ddm(const G& l_grid, double* gx, double* gy, double* gz, const double* psi, double coef, double dx,
                     double dy, double dz) {
    qg_grid* g = cached_grid(l_grid.SizeX(), l_grid.SizeY(), l_grid.SizeZ(), dx, dy, dz);
    const size_t ng = (size_t)l_grid.Size3D();
    Staged<double> sx(gx, ng, false), sy(gy, ng, false), sz(gz, ng, false);
    Staged<double> sp(const_cast<double*>(psi), ng, true);
    check(qg_gradient8_apply(g, QG_F64, 1, sp.get(), (int64_t)ng, sx.get(), sy.get(), sz.get(), (int64_t)ng, coef,
                             nullptr),
          "qg_gradient8_apply");
    sx.sync_back();
    sy.sync_back();
    sz.sync_back();
}

// Host-array form kept for existing callers of this header.
template <class T, class G>
void Laplasian8thHost(const G& g, T* out, const T* p, double coef, double dx, double dy, double dz) {
    Laplasian8th<T>(g, out, p, coef, dx, dy, dz);
}

// Nonlocal KB projectors for one process holding the whole grid (PseudoPotIntegrator_mpi).
// Load() takes the projector channels of an element directly (qg_proj_spec, Gaussian or log-grid
// table radial part); Load(Z, filepath) reads them from an MBK .vps file (qg_vps_*).
class PseudoPotIntegratorGPU {
public:
    ~PseudoPotIntegratorGPU() {
        if (nl_) qg_nonlocal_destroy(nl_);
    }
    // InitializeGrid(grid, dx, dy, dz) (PseudoPotOperator.h:93)
    template <class G>
    void InitializeGrid(const G& g, double dx, double dy, double dz) {
        nx_ = g.SizeX(); ny_ = g.SizeY(); nz_ = g.SizeZ();
        d_[0] = dx; d_[1] = dy; d_[2] = dz;
        grid_.reset(new Grid(nx_, ny_, nz_, dx, dy, dz));
    }
    void Load(int Z, const std::vector<qg_proj_spec>& projectors) {
        if (nl_) throw std::runtime_error("Load before the first UpdatePosition");
        for (auto& e : elements_)
            if (e.first == Z) { e.second = projectors; return; }
        elements_.emplace_back(Z, projectors);
    }
    // Load(Z, filepath, comm) (PseudoPotOperator_vlocal.h:80-91): the element's projector
    // channels from an MBK .vps file; a second Load of the same Z keeps the first, as the
    // reference's m_pp_list.emplace does.  The communicator is not used (one process).
    void Load(int Z, const char* filepath) {
        if (nl_) throw std::runtime_error("Load before the first UpdatePosition");
        for (auto& e : elements_)
            if (e.first == Z) return;
        qg_vps* v = nullptr;
        check(qg_vps_load(filepath, &v), "qg_vps_load");
        qg_vps_info info{};
        std::vector<qg_proj_spec> specs;
        int rc = qg_vps_get_info(v, &info);
        if (rc == QG_OK) {
            specs.resize(std::max(1, info.num_projectors));
            rc = qg_vps_projector_specs(v, specs.data(), (int)specs.size());
            specs.resize(info.num_projectors);
        }
        if (rc == QG_OK) {
            // the tables live in v: keep copies owned here
            for (auto& s : specs) {
                tables_.emplace_back(s.table, s.table + s.n_table);
                s.table = tables_.back().data();
            }
        }
        qg_vps_destroy(v);
        check(rc, "qg_vps_projector_specs");
        elements_.emplace_back(Z, specs);
    }
    template <class Comm>
    void Load(int Z, const char* filepath, Comm&) {
        Load(Z, filepath);
    }
    // UpdatePosition(nuclei, num_nuclei, l_grid) (PseudoPotOperator_mpi_v2.h:231-248): the first
    // call builds the projector geometry, later calls move the nuclei of the same handle (every
    // HamiltonianGPU built on this integrator follows).
    template <class G>
    void UpdatePosition(const Nucleus* nuclei, int num_nuclei, const G& l_grid) {
        if (!grid_) throw std::runtime_error("InitializeGrid first");
        if (l_grid.SizeX() != nx_ || l_grid.SizeY() != ny_ || l_grid.SizeZ() != nz_)
            throw std::runtime_error("UpdatePosition: grid differs from InitializeGrid");
        UpdatePosition(nuclei, num_nuclei);
    }
    void UpdatePosition(const Nucleus* nuclei, int n) {
        if (!grid_) throw std::runtime_error("InitializeGrid first");
        std::vector<double> xyz(3 * (size_t)n);
        std::vector<int> species(n);
        for (int a = 0; a < n; ++a) {
            xyz[3 * a] = nuclei[a].Rx; xyz[3 * a + 1] = nuclei[a].Ry; xyz[3 * a + 2] = nuclei[a].Rz;
            species[a] = -1;
            for (size_t s = 0; s < elements_.size(); ++s)
                if (elements_[s].first == nuclei[a].Z) species[a] = (int)s;
            if (species[a] < 0) throw std::runtime_error("ERROR: VPS is not loaded: Z = " + std::to_string(nuclei[a].Z));
        }
        if (nl_) {
            check(qg_nonlocal_update(nl_, n, xyz.data(), species.data()), "qg_nonlocal_update");
        } else {
            std::vector<int> nproj;
            std::vector<qg_proj_spec> specs;
            for (auto& e : elements_) {
                nproj.push_back((int)e.second.size());
                specs.insert(specs.end(), e.second.begin(), e.second.end());
            }
            check(qg_nonlocal_create(grid_->get(), n, xyz.data(), species.data(), (int)nproj.size(), nproj.data(),
                                     specs.data(), &nl_),
                  "qg_nonlocal_create");
        }
        natoms_ = n;
        species_ = species;
    }
    // ProjectionPP(Hp, psi, l_grid, nuclei, num_nuclei) (PseudoPotOperator_mpi_v2.h:465):
    // Hp += V_NL psi for one band, with the nuclei of the last UpdatePosition.
    template <class G>
    void ProjectionPP(double* Hp, const double* psi, const G& l_grid, const Nucleus* nuclei, int num_nuclei) {
        (void)nuclei;
        if ((size_t)l_grid.Size3D() != size())
            throw std::runtime_error("ProjectionPP: grid differs from InitializeGrid");
        if (num_nuclei != natoms_) throw std::runtime_error("ProjectionPP: nuclei differ from UpdatePosition");
        ProjectionPP_bundle(1, Hp, psi);
    }
    // Band-block form on band-major blocks (the engine reads band-major blocks directly, so the
    // reference's transposes around ProjectionPP_bundle, qumasun_mpi.matrix.h:60-66, are not
    // needed).
    void ProjectionPP_bundle(int nbands, double* Hp, const double* psi) {
        if (!nl_) throw std::runtime_error("UpdatePosition first");
        const size_t ng = size();
        Staged<double> sh(Hp, ng * nbands, true);
        Staged<double> sp(const_cast<double*>(psi), ng * nbands, true);
        check(qg_nonlocal_apply(nl_, QG_F64, nbands, sp.get(), (int64_t)ng, sh.get(), (int64_t)ng, 1, nullptr),
              "qg_nonlocal_apply");
        hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
        sh.sync_back();
    }
    // EnergyPPnonlocal(psi, l_grid, nuclei, num_nuclei) (PseudoPotOperator_mpi_v2.h:573-644):
    // sum_j E_j c_j^2 of one band, c = dV sum beta psi
    template <class G>
    double EnergyPPnonlocal(const double* psi, const G& l_grid, const Nucleus* nuclei, int num_nuclei) {
        (void)nuclei;
        check_call(l_grid, num_nuclei, "EnergyPPnonlocal");
        Staged<double> sp(const_cast<double*>(psi), size(), true);
        const double occ = 1.0;
        double e = 0.0;
        check(qg_nonlocal_energy_force(nl_, QG_F64, 1, sp.get(), (int64_t)size(), &occ, &e, nullptr, nullptr),
              "qg_nonlocal_energy_force");
        return e;
    }
    // ForceNonlocal(forces, psi, dpsi_dx, l_grid, nuclei, num_nuclei) (:905-1013): forces[3a + d]
    // += sum_j E_j c_j d_j with c = dV sum beta psi and d = dV sum beta dpsi_d, where dpsi_dx holds
    // the x, y and z gradient blocks one after another, scaled by the caller (-2 occ in
    // mGetForcePseudoNonlocal, qumasun_mpi.force.h:226-258).  Returns sum_j E_j c_j^2.
    template <class G>
    double ForceNonlocal(double* forces, const double* psi, const double* dpsi_dx, const G& l_grid,
                         const Nucleus* nuclei, int num_nuclei) {
        (void)nuclei;
        check_call(l_grid, num_nuclei, "ForceNonlocal");
        const size_t ng = size();
        const int ncoef = qg_nonlocal_num_coef(nl_);
        Staged<double> sp(const_cast<double*>(psi), ng, true);
        Staged<double> sd(const_cast<double*>(dpsi_dx), 3 * ng, true);
        double* cdev = nullptr;
        hip_check(hipMalloc(&cdev, 4 * (size_t)ncoef * sizeof(double)), "hipMalloc");
        int rc = qg_nonlocal_project(nl_, QG_F64, 1, sp.get(), (int64_t)ng, cdev, nullptr);
        if (rc == QG_OK) rc = qg_nonlocal_project(nl_, QG_F64, 3, sd.get(), (int64_t)ng, cdev + ncoef, nullptr);
        std::vector<double> c(4 * (size_t)ncoef);
        if (rc == QG_OK)
            rc = hipMemcpy(c.data(), cdev, c.size() * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess ? QG_OK
                                                                                                          : QG_ERR_HIP;
        (void)hipFree(cdev);
        check(rc, "ForceNonlocal");
        // coefficients in the order atom, projector, m; each projector's KB energy
        double e = 0.0;
        int j = 0;
        for (int a = 0; a < natoms_; ++a)
            for (const qg_proj_spec& sp_ : elements_[species_[a]].second)
                for (int m = 0; m < 2 * sp_.l + 1; ++m, ++j) {
                    e += sp_.energy * c[j] * c[j];
                    for (int d = 0; d < 3; ++d) forces[3 * a + d] += sp_.energy * c[j] * c[(size_t)(d + 1) * ncoef + j];
                }
        return e;
    }
    // mGetEnergyPseudoNonlocal + mGetForcePseudoNonlocal in one call on a band block (the
    // gradients are formed on the device): *energy = sum over bands with occ > 0 of
    // occ sum E c^2, forces (3 per atom, may be NULL) set, not accumulated.
    void EnergyForceNonlocal(int nbands, const double* psi, const double* occupancy, double* energy, double* forces) {
        if (!nl_) throw std::runtime_error("UpdatePosition first");
        Staged<double> sp(const_cast<double*>(psi), size() * nbands, true);
        check(qg_nonlocal_energy_force(nl_, QG_F64, nbands, sp.get(), (int64_t)size(), occupancy, energy, forces,
                                       nullptr),
              "qg_nonlocal_energy_force");
    }
    qg_nonlocal* handle() const { return nl_; }
    qg_grid* grid_handle() const { return grid_->get(); }
    size_t size() const { return (size_t)nx_ * ny_ * nz_; }
    double dvol() const { return d_[0] * d_[1] * d_[2]; }

private:
    template <class G>
    void check_call(const G& l_grid, int num_nuclei, const char* what) const {
        if (!nl_) throw std::runtime_error(std::string(what) + ": UpdatePosition first");
        if ((size_t)l_grid.Size3D() != size()) throw std::runtime_error(std::string(what) + ": grid differs");
        if (num_nuclei != natoms_) throw std::runtime_error(std::string(what) + ": nuclei differ from UpdatePosition");
    }
    std::vector<int> species_;
    int nx_ = 0, ny_ = 0, nz_ = 0, natoms_ = 0;
    double d_[3] = {0, 0, 0};
    std::unique_ptr<Grid> grid_;
    std::vector<std::pair<int, std::vector<qg_proj_spec>>> elements_;
    std::deque<std::vector<double>> tables_;   // radial tables of the .vps channels
    qg_nonlocal* nl_ = nullptr;
};

// H = V + V_NL - 1/2 Lap8 (QUMASUN_MPI::mHamiltonianMatrix_ddm) for band blocks.  As OpeA of
// Eigen_d_multi_mpi it is applied on the device to the whole block at once.
class HamiltonianGPU {
public:
    HamiltonianGPU(PseudoPotIntegratorGPU& pp, const double* Vtot_host) : pp_(pp) {
        check(qg_hamiltonian_create(pp.grid_handle(), Vtot_host, pp.handle(), &h_), "qg_hamiltonian_create");
    }
    ~HamiltonianGPU() { qg_hamiltonian_destroy(h_); }
    HamiltonianGPU(const HamiltonianGPU&) = delete;
    HamiltonianGPU& operator=(const HamiltonianGPU&) = delete;
    void SetPotential(const double* Vtot_host) { check(qg_hamiltonian_set_potential(h_, Vtot_host), "set_potential"); }
    // Hp = H p for nbands band-major columns (FP64 or FP32, host or device)
    template <class T>
    void Apply(int nbands, T* Hp, const T* p, hipStream_t stream = nullptr) {
        const size_t ng = pp_.size();
        Staged<T> sh(Hp, ng * nbands, false);
        Staged<T> sp(const_cast<T*>(p), ng * nbands, true);
        check(qg_hamiltonian_apply(h_, dtype_of<T>::value, nbands, sp.get(), (int64_t)ng, sh.get(), (int64_t)ng,
                                   stream),
              "qg_hamiltonian_apply");
        if (sh.host()) {
            hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            sh.sync_back();
        }
    }
    // OpeA(Ax, x) of the reference (one band)
    void operator()(double* Ax, double* x) { Apply<double>(1, Ax, x); }
    qg_ham* get() const { return h_; }
    size_t size() const { return pp_.size(); }
    double dvol() const { return pp_.dvol(); }

private:
    PseudoPotIntegratorGPU& pp_;
    qg_ham* h_ = nullptr;
};

// The k-point operator of QUMASUN_KPOINT (qumasun_kpoint.matrix.h:21-110) on one GPU, for one
// spin-k set: H_k of a HamiltonianGPU at k = (kx, ky, kz) bohr^-1 (kpoint_i 2 pi / (L_i samples_i)
// in the reference).  Complex bands in the reference's layout: per band ng real parts, then ng
// imaginary parts (SoAComplex with re, im pointers), converted on the device to the engine's
// interleaved complex.
class KpointHamiltonianGPU {
public:
    KpointHamiltonianGPU(HamiltonianGPU& H, double kx, double ky, double kz) : H_(H) {
        check(qg_kpoint_create(H.get(), kx, ky, kz, &kp_), "qg_kpoint_create");
    }
    ~KpointHamiltonianGPU() {
        qg_kpoint_destroy(kp_);
        if (zin_) (void)hipFree(zin_);
        if (zout_) (void)hipFree(zout_);
    }
    KpointHamiltonianGPU(const KpointHamiltonianGPU&) = delete;
    KpointHamiltonianGPU& operator=(const KpointHamiltonianGPU&) = delete;
    // Ax = H_k x for nbands planar complex bands (2 ng doubles each; host or device)
    void ApplyPlanar(int nbands, double* Ax, const double* x, hipStream_t stream = nullptr) {
        const size_t ng = H_.size();
        Staged<double> sa(Ax, 2 * ng * nbands, false);
        Staged<double> sx(const_cast<double*>(x), 2 * ng * nbands, true);
        reserve(nbands);
        check(qg_complex_from_planar((int64_t)ng, nbands, sx.get(), QG_Z128, zin_, stream), "qg_complex_from_planar");
        check(qg_kpoint_apply(kp_, QG_Z128, nbands, zin_, (int64_t)ng, zout_, (int64_t)ng, stream), "qg_kpoint_apply");
        check(qg_complex_to_planar((int64_t)ng, nbands, zout_, QG_Z128, sa.get(), stream), "qg_complex_to_planar");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        sa.sync_back();
    }
    // OpeA(Ax, x) of the reference's k-point solver: one band with re / im host arrays of ng
    template <class SoA>
    void operator()(SoA& Ax, const SoA& x) {
        const size_t ng = H_.size();
        std::vector<double> in(2 * ng), out(2 * ng);
        std::copy(x.re, x.re + ng, in.begin());
        std::copy(x.im, x.im + ng, in.begin() + ng);
        ApplyPlanar(1, out.data(), in.data());
        std::copy(out.begin(), out.begin() + ng, Ax.re);
        std::copy(out.begin() + ng, out.end(), Ax.im);
    }
    qg_kpoint* get() const { return kp_; }
    size_t size() const { return H_.size(); }

private:
    void reserve(int nbands) {
        const size_t need = H_.size() * nbands * 2 * sizeof(double);
        if (need <= cap_) return;
        if (zin_) (void)hipFree(zin_);
        if (zout_) (void)hipFree(zout_);
        hip_check(hipMalloc(&zin_, need), "hipMalloc");
        hip_check(hipMalloc(&zout_, need), "hipMalloc");
        cap_ = need;
    }
    HamiltonianGPU& H_;
    qg_kpoint* kp_ = nullptr;
    void* zin_ = nullptr;
    void* zout_ = nullptr;
    size_t cap_ = 0;
};

// GramSchmidt(N, num_solution, psi_set, dVol) (GramSchmidt.h:6-38) / GramSchmidt_ddm
// (GramSchmidt_mpi.h:9-53) on a num_solution x N FP64 band-major block, in place.
inline void GramSchmidt(int64_t N, int num_solution, double* psi, double dVol) {
    Staged<double> sp(psi, (size_t)N * num_solution, true);
    check(qg_gram_schmidt(QG_F64, N, num_solution, sp.get(), N, dVol, nullptr), "qg_gram_schmidt");
    sp.sync_back();
}

namespace detail {

template <class WATCH>
void watch_thunk(void* user, int sector, int64_t count) {
    static_cast<WATCH*>(user)->Record(sector, count);
}

// A user operator OpeA(double* Ax, double* x) for one band, called from the engine's operator
// callback per band, on pointers in the caller's memory space.
template <class OperationA>
struct UserOp {
    OperationA* op;
    size_t ng;
    bool host;                    // the caller's arrays are host memory: stage each block
    std::vector<double> hx, hax;
    static void call(void* user, int dtype, int nbands, const void* x, void* ax, void* stream) {
        UserOp* c = static_cast<UserOp*>(user);
        (void)dtype;   // the compat solvers run in FP64
        hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
        const size_t n = c->ng * (size_t)nbands;
        if (c->host) {
            c->hx.resize(n);
            c->hax.resize(n);
            hip_check(hipMemcpy(c->hx.data(), x, n * 8, hipMemcpyDeviceToHost), "hipMemcpy");
            for (int k = 0; k < nbands; ++k) (*c->op)(c->hax.data() + k * c->ng, c->hx.data() + k * c->ng);
            hip_check(hipMemcpy(ax, c->hax.data(), n * 8, hipMemcpyHostToDevice), "hipMemcpy");
        } else {
            for (int k = 0; k < nbands; ++k)
                (*c->op)(static_cast<double*>(ax) + k * c->ng,
                         const_cast<double*>(static_cast<const double*>(x)) + k * c->ng);
            hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        }
    }
};

// A Hamiltonian of V = 0 without projectors on an (nx, ny, nz) grid of cell dVol: the grid the
// engine's solver needs when OpeA is a foreign operator.
struct GridOnlyHam {
    std::unique_ptr<Grid> g;
    qg_ham* h = nullptr;
    GridOnlyHam(int nx, int ny, int nz, double dvol) {
        const double d = std::cbrt(dvol);
        g.reset(new Grid(nx, ny, nz, d, d, d));
        std::vector<double> v((size_t)nx * ny * nz, 0.0);
        check(qg_hamiltonian_create(g->get(), v.data(), nullptr, &h), "qg_hamiltonian_create");
    }
    ~GridOnlyHam() { qg_hamiltonian_destroy(h); }
};

template <class OperationA>
inline constexpr bool is_engine_op = std::is_same<std::decay_t<OperationA>, HamiltonianGPU*>::value ||
                                     std::is_same<std::decay_t<OperationA>, std::reference_wrapper<HamiltonianGPU>>::value;

inline HamiltonianGPU& engine_of(HamiltonianGPU* h) { return *h; }
inline HamiltonianGPU& engine_of(std::reference_wrapper<HamiltonianGPU> h) { return h.get(); }

// One solve: eigen_vectors / keep_P of num_solution x ng (host or device), OpeA the engine's
// Hamiltonian (applied to whole blocks on the device) or a per-band user operator.
template <class OperationA, class WATCH>
void solve(int nx, int ny, int nz, double dVol, int num_solution, int iter_max, double* eigen_values,
           double* eigen_vectors, double* keep_P, int SCF_step, double tol, OperationA OpeA, WATCH& watch) {
    const size_t ng = (size_t)nx * ny * nz;
    qg_lobpcg_opts o;
    qg_lobpcg_default_opts(&o);
    o.max_iter = iter_max;
    o.tol = tol;
    o.precision = QG_PREC_F64;
    o.watch = &watch_thunk<WATCH>;
    o.watch_user = &watch;
    Staged<double> sx(eigen_vectors, ng * num_solution, true);
    Staged<double> sp(keep_P, ng * num_solution, SCF_step != 0);
    qg_lobpcg_stats st{};
    if constexpr (is_engine_op<OperationA>) {
        HamiltonianGPU& H = engine_of(OpeA);
        if (H.size() != ng) throw std::runtime_error("Eigen_d_multi_mpi: grid differs from the Hamiltonian's");
        if (std::fabs(H.dvol() - dVol) > 1e-12 * dVol)
            throw std::runtime_error("Eigen_d_multi_mpi: dVol differs from the Hamiltonian's grid");
        check(qg_lobpcg_d_multi(H.get(), &o, num_solution, eigen_values, sx.get(), sp.get(), SCF_step, &st),
              "qg_lobpcg_d_multi");
    } else {
        GridOnlyHam gh(nx, ny, nz, dVol);
        UserOp<OperationA> u{&OpeA, ng, sx.host(), {}, {}};
        o.op = &UserOp<OperationA>::call;
        o.op_user = &u;
        check(qg_lobpcg_d_multi(gh.h, &o, num_solution, eigen_values, sx.get(), sp.get(), SCF_step, &st),
              "qg_lobpcg_d_multi");
    }
    sx.sync_back();
    sp.sync_back();
}

}  // namespace detail

namespace LOBPCG {
// Eigen_z_multi_mpi's task (lobpcg_z_multi_mpi.h:988) on the device for one spin-k set: the
// lowest num_solution eigenpairs of H_k by the engine's complex FP64 LOBPCG, from the planar
// complex start X (num_solution bands of 2 ng doubles, host or device), overwritten with the
// eigenvectors.  Stops at tol (largest residual norm) or iter_max iterations; returns the
// iterations run.  precond_shift > 0 preconditions with (T_k + shift)^-1 (not in the reference).
inline int Eigen_z_multi(KpointHamiltonianGPU& Hk, int num_solution, int iter_max, double* eigen_values, double* X,
                         double tol, double* max_residual = nullptr, hipStream_t stream = nullptr,
                         double precond_shift = 0.0) {
    const size_t ng = Hk.size();
    Staged<double> sx(X, 2 * ng * num_solution, true);
    void* z = nullptr;
    hip_check(hipMalloc(&z, 2 * ng * num_solution * sizeof(double)), "hipMalloc");
    int iters = 0;
    double res = 0.0;
    int rc = qg_complex_from_planar((int64_t)ng, num_solution, sx.get(), QG_Z128, z, stream);
    if (rc == QG_OK)
        rc = qg_kpoint_lobpcg(Hk.get(), num_solution, z, (int64_t)ng, eigen_values, iter_max, tol, 0, precond_shift,
                              &iters, &res, stream);
    if (rc == QG_OK) rc = qg_complex_to_planar((int64_t)ng, num_solution, z, QG_Z128, sx.get(), stream);
    if (rc == QG_OK) rc = hipStreamSynchronize(stream) == hipSuccess ? QG_OK : QG_ERR_HIP;
    (void)hipFree(z);
    check(rc, "Eigen_z_multi");
    sx.sync_back();
    if (max_residual) *max_residual = res;
    return iters;
}


// Eigen_d_multi_mpi(l_grid, dVol, num_solution, iter_max, eigen_values, eigen_vectors, keep_P,
// work, keep_S_matrix, SCF_step, OpeA, watch) (lobpcg_d_multi_mpi.h:882-890): the reference's
// block LOBPCG for iter_max iterations (2N basis on SCF_step 0, refresh every 10 iterations,
// rollback at |1 - ||x'||^2| > 1e-3), on the device in FP64.
//   OpeA: a HamiltonianGPU* (or std::ref of one), applied to whole blocks on the device, or any
//         callable OpeA(double* Ax, double* x) for one band (the per-band lambda of
//         QUMASUN_MPI::mSolveLOBPCG, qumasun_mpi.lobpcg.h:33-35), given pointers in the memory
//         space of eigen_vectors.
//   work, keep_S_matrix: accepted for the reference's signature; the engine keeps its work
//         blocks and Gram state in the Hamiltonian's solver state (the reference reads
//         keep_S_matrix only under S_MATRIX_RELOAD, lobpcg_d_multi_mpi.h:46, which it leaves off).
//   watch: WATCH::Record(sector, count) with the reference's sector ids (10, 11, 15, 16, 17).
template <class G, class OperationA, class WATCH>
void Eigen_d_multi_mpi(const G& l_grid, double dVol, int num_solution, int iter_max, double* eigen_values,
                       double* eigen_vectors, double* keep_P, double* work, double* keep_S_matrix, int SCF_step,
                       OperationA OpeA, WATCH& watch) {
    (void)work;
    (void)keep_S_matrix;
    detail::solve(l_grid.SizeX(), l_grid.SizeY(), l_grid.SizeZ(), dVol, num_solution, iter_max, eigen_values,
                  eigen_vectors, keep_P, SCF_step, 0.0, OpeA, watch);
}

// Work size of Eigen_d_multi_mpi in doubles (lobpcg_d_multi_mpi.h:1575-1577).
inline size_t WorkSize_d_multi_mpi(int local_size, int num_solution) { return 5 * (size_t)local_size * num_solution; }

// Earlier engine-shaped forms (a Hamiltonian and device blocks), kept for existing callers.
inline qg_lobpcg_stats Eigen_d_multi(HamiltonianGPU& H, int num_solution, int iter_max, double* eigen_values,
                                     double* eigen_vectors_dev, double* keep_P_dev, int SCF_step,
                                     const qg_lobpcg_opts* opts = nullptr) {
    qg_lobpcg_opts o;
    qg_lobpcg_default_opts(&o);
    if (opts) o = *opts;
    o.max_iter = iter_max;
    qg_lobpcg_stats st{};
    check(qg_lobpcg_d_multi(H.get(), &o, num_solution, eigen_values, eigen_vectors_dev, keep_P_dev, SCF_step, &st),
          "qg_lobpcg_d_multi");
    return st;
}

}  // namespace LOBPCG

// EigenLOBPCG_d_multi(N, dVol, num_solution, iter_max, eigen_values, eigen_vectors, keep, work,
// is_first_time, OpeA, watch) (lobpcg_d_multi.h:23-27), the reference's single-process form:
// N grid points (a cube of side N^(1/3) is assumed when OpeA is a user operator; with a
// HamiltonianGPU the Hamiltonian's grid is used), keep = the conjugate block, is_first_time
// = SCF step 0.  Stops as the reference does when every band's dVol-weighted squared residual
// is below 1e-10 dVol (lobpcg_d_multi.h:59, :533), or after iter_max iterations.
template <class OperationA, class WATCH>
void EigenLOBPCG_d_multi(int N, double dVol, int num_solution, int iter_max, double* eigen_values,
                         double* eigen_vectors, double* keep, double* work, bool is_first_time, OperationA OpeA,
                         WATCH& watch) {
    (void)work;
    int nx, ny, nz;
    if constexpr (detail::is_engine_op<OperationA>) {
        (void)N;
        nx = (int)detail::engine_of(OpeA).size();
        ny = nz = 1;
    } else {
        nx = (int)std::lround(std::cbrt((double)N));
        if ((int64_t)nx * nx * nx != N) throw std::runtime_error("EigenLOBPCG_d_multi: N is not a cube; use Eigen_d_multi_mpi");
        ny = nz = nx;
    }
    // max_k dVol sum r_k^2 < 1e-10 dVol  <=>  max_k ||r_k|| (dVol-normalised) < sqrt(1e-10 dVol)
    const double tol = std::sqrt(1.0e-10 * dVol);
    if constexpr (detail::is_engine_op<OperationA>) {
        HamiltonianGPU& H = detail::engine_of(OpeA);
        qg_lobpcg_opts o;
        qg_lobpcg_default_opts(&o);
        o.max_iter = iter_max;
        o.tol = tol;
        o.watch = &detail::watch_thunk<WATCH>;
        o.watch_user = &watch;
        const size_t ng = H.size();
        Staged<double> sx(eigen_vectors, ng * num_solution, true);
        Staged<double> sp(keep, ng * num_solution, !is_first_time);
        qg_lobpcg_stats st{};
        check(qg_lobpcg_d_multi(H.get(), &o, num_solution, eigen_values, sx.get(), sp.get(), is_first_time ? 0 : 1,
                                &st),
              "qg_lobpcg_d_multi");
        sx.sync_back();
        sp.sync_back();
    } else {
        detail::solve(nx, ny, nz, dVol, num_solution, iter_max, eigen_values, eigen_vectors, keep,
                      is_first_time ? 0 : 1, tol, OpeA, watch);
    }
}

// Work size of EigenLOBPCG_d_multi in doubles (lobpcg_d_multi.h:548-551).
inline size_t WorkSizeLOBPCG_d_multi(int line_size, int num_solution) { return 5 * (size_t)line_size * num_solution; }

}  // namespace qumasun_gpu
