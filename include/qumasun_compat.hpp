// qumasun_compat.hpp — C++ call surface of QUMASUN's H·psi + LOBPCG path on top of the C ABI
// (qumasun_gpu.h), so the reference's call sites can switch to the MI355X engine with the
// same argument meaning.  Header-only; errors become std::runtime_error (the reference
// returns void and prints; the engine reports every failure).
//
//   Laplasian8th<T>(GridRange, out, p, coef, dx, dy, dz)   GridDifference8th.h:177-185
//   PseudoPotIntegratorGPU::UpdatePosition / ProjectionPP   PseudoPotOperator_mpi_v2.h:231, :465
//   PseudoPotIntegratorGPU::ProjectionPP_bundle             PseudoPotOperator_mpi_v2.h:789
//   HamiltonianGPU::Apply  (mHamiltonianMatrix_ddm)         qumasun_mpi.matrix.h:21-43
//   LOBPCG::Eigen_d_multi                                   lobpcg_d_multi_mpi.h:882-890
//
// Pointers named *_dev are device memory; the *Host overloads stage host arrays through
// device memory (one copy each way) for drop-in use from host-resident code.
#pragma once
#include <hip/hip_runtime.h>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include "qumasun_gpu.h"

namespace qumasun_gpu {

inline void check(int rc, const char* what) {
    if (rc != QG_OK) throw std::runtime_error(std::string(what) + ": " + qg_last_error());
}
inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// The reference's box of grid indices [begin, end) (GridRange.h:14-46); one process holds the
// whole periodic grid, so begin_* = 0.
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

// RAII grid handle
class Grid {
public:
    Grid(const GridRange& g, double dx, double dy, double dz) {
        check(qg_grid_create(g.SizeX(), g.SizeY(), g.SizeZ(), dx, dy, dz, &h_), "qg_grid_create");
    }
    ~Grid() { qg_grid_destroy(h_); }
    Grid(const Grid&) = delete;
    Grid& operator=(const Grid&) = delete;
    qg_grid* get() const { return h_; }

private:
    qg_grid* h_ = nullptr;
};

// out += coef * Lap8 p (periodic), one band, device pointers — Laplasian8th<T>.
template <class T>
void Laplasian8th(const GridRange& g, T* out_dev, const T* p_dev, double coef, double dx, double dy, double dz,
                  hipStream_t stream = nullptr) {
    Grid grid(g, dx, dy, dz);
    const int64_t ng = (int64_t)g.Size3D();
    check(qg_laplacian8_apply(grid.get(), dtype_of<T>::value, 1, p_dev, ng, out_dev, ng, coef, 1, stream),
          "qg_laplacian8_apply");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

// Host-pointer overload (stages through device memory).
template <class T>
void Laplasian8thHost(const GridRange& g, T* out, const T* p, double coef, double dx, double dy, double dz) {
    const size_t bytes = g.Size3D() * sizeof(T);
    T *dp = nullptr, *dout = nullptr;
    hip_check(hipMalloc(&dp, bytes), "hipMalloc");
    hip_check(hipMalloc(&dout, bytes), "hipMalloc");
    hip_check(hipMemcpy(dp, p, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(dout, out, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    Laplasian8th<T>(g, dout, dp, coef, dx, dy, dz);
    hip_check(hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    (void)hipFree(dp);
    (void)hipFree(dout);
}

// Nonlocal KB projectors for one process holding the whole grid (PseudoPotIntegrator_mpi).
// Load() takes the projector channels of an element directly (the .vps loader is not part of
// this engine): qg_proj_spec with a Gaussian or a log-grid table radial part.
class PseudoPotIntegratorGPU {
public:
    ~PseudoPotIntegratorGPU() { release(); }
    void InitializeGrid(const GridRange& g, double dx, double dy, double dz) {
        g_ = g;
        d_[0] = dx; d_[1] = dy; d_[2] = dz;
        grid_.reset(new Grid(g, dx, dy, dz));
    }
    void Load(int Z, const std::vector<qg_proj_spec>& projectors) {
        for (auto& e : elements_)
            if (e.first == Z) { e.second = projectors; return; }
        elements_.emplace_back(Z, projectors);
    }
    // UpdatePosition(nuclei, n, l_grid) (PseudoPotOperator_mpi_v2.h:231-248)
    void UpdatePosition(const Nucleus* nuclei, int n) {
        if (!grid_) throw std::runtime_error("InitializeGrid first");
        std::vector<double> xyz(3 * (size_t)n);
        std::vector<int> species(n), nproj;
        std::vector<qg_proj_spec> specs;
        for (auto& e : elements_) {
            nproj.push_back((int)e.second.size());
            specs.insert(specs.end(), e.second.begin(), e.second.end());
        }
        for (int a = 0; a < n; ++a) {
            xyz[3 * a] = nuclei[a].Rx; xyz[3 * a + 1] = nuclei[a].Ry; xyz[3 * a + 2] = nuclei[a].Rz;
            species[a] = -1;
            for (size_t s = 0; s < elements_.size(); ++s)
                if (elements_[s].first == nuclei[a].Z) species[a] = (int)s;
            if (species[a] < 0) throw std::runtime_error("ERROR: VPS is not loaded: Z = " + std::to_string(nuclei[a].Z));
        }
        release();
        check(qg_nonlocal_create(grid_->get(), n, xyz.data(), species.data(), (int)nproj.size(), nproj.data(),
                                 specs.data(), &nl_),
              "qg_nonlocal_create");
    }
    // Hp += V_NL psi for one band (device pointers) — ProjectionPP.
    void ProjectionPP(double* Hp_dev, const double* psi_dev, hipStream_t stream = nullptr) {
        ProjectionPP_bundle(1, Hp_dev, psi_dev, stream);
    }
    // Band-block form; band-major layout (the engine reads band-major blocks directly, so the
    // reference's transposes around ProjectionPP_bundle are not needed).
    void ProjectionPP_bundle(int nbands, double* Hp_dev, const double* psi_dev, hipStream_t stream = nullptr) {
        const int64_t ng = (int64_t)g_.Size3D();
        check(qg_nonlocal_apply(nl_, QG_F64, nbands, psi_dev, ng, Hp_dev, ng, 1, stream), "qg_nonlocal_apply");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    qg_nonlocal* handle() const { return nl_; }
    const GridRange& grid() const { return g_; }
    qg_grid* grid_handle() const { return grid_->get(); }

private:
    void release() {
        if (nl_) qg_nonlocal_destroy(nl_);
        nl_ = nullptr;
    }
    GridRange g_{};
    double d_[3] = {0, 0, 0};
    std::unique_ptr<Grid> grid_;
    std::vector<std::pair<int, std::vector<qg_proj_spec>>> elements_;
    qg_nonlocal* nl_ = nullptr;
};

// H = V + V_NL - 1/2 Lap8 (QUMASUN_MPI::mHamiltonianMatrix_ddm) for band blocks.
class HamiltonianGPU {
public:
    HamiltonianGPU(PseudoPotIntegratorGPU& pp, const double* Vtot_host) : pp_(pp) {
        check(qg_hamiltonian_create(pp.grid_handle(), Vtot_host, pp.handle(), &h_), "qg_hamiltonian_create");
    }
    ~HamiltonianGPU() { qg_hamiltonian_destroy(h_); }
    HamiltonianGPU(const HamiltonianGPU&) = delete;
    HamiltonianGPU& operator=(const HamiltonianGPU&) = delete;
    void SetPotential(const double* Vtot_host) { check(qg_hamiltonian_set_potential(h_, Vtot_host), "set_potential"); }
    // Hp = H p for nbands band-major device columns (FP64 or FP32)
    template <class T>
    void Apply(int nbands, T* Hp_dev, const T* p_dev, hipStream_t stream = nullptr) {
        const int64_t ng = (int64_t)pp_.grid().Size3D();
        check(qg_hamiltonian_apply(h_, dtype_of<T>::value, nbands, p_dev, ng, Hp_dev, ng, stream),
              "qg_hamiltonian_apply");
    }
    qg_ham* get() const { return h_; }

private:
    PseudoPotIntegratorGPU& pp_;
    qg_ham* h_ = nullptr;
};

// GramSchmidt(N, num_solution, psi_set, dVol) (GramSchmidt.h:6-38) / GramSchmidt_ddm
// (GramSchmidt_mpi.h:9-53) on a num_solution x N FP64 band-major device block, in place.
inline void GramSchmidt(int64_t N, int num_solution, double* psi_dev, double dVol) {
    check(qg_gram_schmidt(QG_F64, N, num_solution, psi_dev, N, dVol, nullptr), "qg_gram_schmidt");
}

namespace LOBPCG {

// Eigen_d_multi_mpi(l_grid, dVol, num_solution, iter_max, eigen_values, eigen_vectors, keep_P,
// work, keep_S_matrix, SCF_step, OpeA, watch) with OpeA = the device Hamiltonian.  dVol is
// taken from the grid; the work block is owned by the engine; eigen_vectors_dev and
// keep_P_dev are num_solution x Ng FP64 device blocks.  Options default to the reference's
// behaviour (FP64, fixed iter_max, refresh every 10 iterations, rollback at 1e-3).
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

// The reference's form with a user operator: OpeA(double* Ax_dev, double* x_dev) applies A to
// one band (device pointers, FP64), as the per-band lambda of QUMASUN_MPI::mSolveLOBPCG
// (qumasun_mpi.lobpcg.h:33-35); H supplies the grid.  The solve runs in FP64.
template <class OperationA>
inline qg_lobpcg_stats Eigen_d_multi(HamiltonianGPU& H, int num_solution, int iter_max, double* eigen_values,
                                     double* eigen_vectors_dev, double* keep_P_dev, int SCF_step, OperationA OpeA,
                                     int64_t grid_size, const qg_lobpcg_opts* opts = nullptr) {
    struct Ctx {
        OperationA* op;
        int64_t ng;
    } ctx{&OpeA, grid_size};
    qg_lobpcg_opts o;
    qg_lobpcg_default_opts(&o);
    if (opts) o = *opts;
    o.max_iter = iter_max;
    o.precision = QG_PREC_F64;
    o.op_user = &ctx;
    o.op = [](void* user, int dtype, int nbands, const void* x, void* ax, void* stream) {
        Ctx* c = static_cast<Ctx*>(user);
        (void)dtype;
        (void)hipStreamSynchronize((hipStream_t)stream);
        for (int k = 0; k < nbands; ++k)
            (*c->op)(static_cast<double*>(ax) + (int64_t)k * c->ng,
                     const_cast<double*>(static_cast<const double*>(x)) + (int64_t)k * c->ng);
        (void)hipDeviceSynchronize();
    };
    qg_lobpcg_stats st{};
    check(qg_lobpcg_d_multi(H.get(), &o, num_solution, eigen_values, eigen_vectors_dev, keep_P_dev, SCF_step, &st),
          "qg_lobpcg_d_multi");
    return st;
}

}  // namespace LOBPCG
}  // namespace qumasun_gpu
