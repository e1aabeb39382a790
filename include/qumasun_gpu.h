/*
 * qumasun_gpu.h — C ABI of the MI355X Kohn-Sham H·psi + block-LOBPCG engine.
 *
 * Plain C: opaque handles, raw pointers and sizes, int status codes (QG_OK = 0), and
 * qg_last_error() for the message of the last failure on the calling thread.  No C++
 * exceptions cross this boundary and no torch types appear in it.
 *
 * Conventions (same as the reference's real-space grid code):
 *   - grid index i = ix + Nx*(iy + Ny*iz), x fastest (GridDifference8th.h:137-138);
 *   - a block of nbands wavefunctions is band-major: band b starts at psi + b*ld
 *     (lobpcg_d_multi_mpi.h:898-913), ld >= Nx*Ny*Nz;
 *   - grid point positions are i*h with no half-cell offset
 *     (PseudoPotOperator_mpi_v2.h:144-150);
 *   - every psi/hpsi/X/P/work pointer is DEVICE memory; host arrays are named *_host.
 *   - `stream` is a hipStream_t passed as void*; NULL is HIP's null stream.  Calls on one
 *     handle are not thread-safe (neither is the reference's integrator, which holds
 *     mutable per-call state).
 */
#ifndef QUMASUN_GPU_H
#define QUMASUN_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QG_OK 0
#define QG_ERR_ARG 1      /* invalid argument (shape, null pointer, unsupported option) */
#define QG_ERR_HIP 2      /* a HIP runtime call failed */
#define QG_ERR_LIB 3      /* rocBLAS / rocSOLVER failure */
#define QG_ERR_NUMERIC 4  /* numerical breakdown (e.g. Cholesky of a non-SPD Gram) */

/* arithmetic type of a psi block */
#define QG_F64 0
#define QG_F32 1
#define QG_Z128 2   /* complex FP64, re and im interleaved (k-point blocks) */
#define QG_C64 3    /* complex FP32, re and im interleaved (k-point blocks) */

const char* qg_last_error(void);
int qg_version(void);
/* Number of visible GPUs (0 on a host without one); never faults. */
int qg_device_count(void);

/* ------------------------------------------------------------------ grid ------ */
typedef struct qg_grid qg_grid;

/* Periodic Nx*Ny*Nz grid with spacings dx,dy,dz (the reference's GridRange + m_dx/dy/dz,
 * GridRange.h:14-46, qumasun_mpi.constructor.h:35). */
int qg_grid_create(int nx, int ny, int nz, double dx, double dy, double dz, qg_grid** out);
int qg_grid_destroy(qg_grid* g);

/* out[b] = (accumulate ? out[b] : 0) + coef * Lap8 psi[b] for b < nbands, periodic.
 * Replaces Laplasian8th<T>(GridRange, out, p, coef, dx, dy, dz) (GridDifference8th.h:177-185),
 * which always accumulates, and Laplasian8th_ddm (GridDifference8th.h:320-344) with
 * split_dimension 0.  dtype is QG_F64 or QG_F32 (both psi and out). */
int qg_laplacian8_apply(const qg_grid* g, int dtype, int nbands,
                        const void* psi, int64_t ld_psi, void* out, int64_t ld_out,
                        double coef, int accumulate, void* stream);

/* ------------------------------------------------------- nonlocal projectors ---- */
/* One Kleinman-Bylander projector channel of a species: beta_lm(r) = p(r) Y_lm(r/|r|),
 * p(r) = 0 for r^2 > cutoff^2 (PseudoPotOperator_mpi_v2.h:156-166).
 * radial_kind QG_RADIAL_GAUSSIAN: p(r) = amplitude * exp(-r^2/(2 sigma^2)).
 * radial_kind QG_RADIAL_TABLE:  p(r) from a log-grid table (xi = log r, values on
 *   xi_min + k*xi_delta, k < n_table), interpolated as RadialGrid2::GetValueByXi
 *   (RadialGrid2.h:81-136); r^2 < r_min^2 uses table[0]
 *   (PseudoPotOperator_mpi_v2.h:154-158).  The table holds the j-averaged projector. */
#define QG_RADIAL_GAUSSIAN 0
#define QG_RADIAL_TABLE 1
typedef struct qg_proj_spec {
    int l;               /* angular momentum, 0 or 1 */
    double energy;       /* KB energy E_k (Ha), PseudoPot_MBK::projector_energy_up */
    double cutoff;       /* PseudoPot_MBK::cutoff_r[l] (bohr) */
    int radial_kind;
    double sigma;        /* gaussian width (bohr) */
    double amplitude;    /* gaussian prefactor */
    int n_table;         /* table length */
    double xi_min, xi_delta, r_min;
    const double* table; /* host pointer, n_table values */
} qg_proj_spec;

/* ----------------------------------------------------- MBK pseudopotential files ---- */
/* An MBK .vps file read as LoadVPS reads it (vps_loader.h:87-247, file model of
 * JobParam.h:246-384): pseudo.NandL cutoffs per l, project.energies, the Pseudo.Potentials
 * table (xi, r, V_local and two j-components per projector), the optional density.PCC
 * block, valence.electron and local.cutoff; PseudoPot_MBK::SortByL applied as the reference
 * applies it.  Errors (no file, vps.type other than MBK, missing or inconsistent blocks) return
 * QG_ERR_ARG with the message in qg_last_error(). */
typedef struct qg_vps qg_vps;
typedef struct qg_vps_info {
    int valence_electron;      /* (int) valence.electron */
    int num_projectors;
    int num_radial_grids;
    int has_pcc_charge;
    int total_projector_lm;    /* PseudoPot_MBK::TotalProjectorLM */
    int max_projector_l;       /* PseudoPot_MBK::MaxProjectorL */
    double cutoff_vlocal;      /* local.cutoff */
    double cutoff_r[4];        /* largest pseudo.NandL cutoff per l */
    double xi_min, xi_delta;   /* log-grid of the tables */
    double pcc_charge_integral;/* 4 pi xi_delta sum pcc r^3 (the figure LoadVPS prints) */
} qg_vps_info;
int qg_vps_load(const char* path, qg_vps** out);
int qg_vps_destroy(qg_vps* v);
int qg_vps_get_info(const qg_vps* v, qg_vps_info* info);
/* radial grid r, V_local and the PCC charge (zeros without a density.PCC block); each
 * num_radial_grids values, any pointer may be NULL */
int qg_vps_get_radial(const qg_vps* v, double* radius, double* v_local, double* pcc_charge);
/* projector k: l, the two energies and the two j-components (num_radial_grids values each;
 * any pointer may be NULL) */
int qg_vps_get_projector(const qg_vps* v, int k, int* l, double* energy_up, double* energy_down,
                         double* up_values, double* down_values);
/* The projector channels of the file for qg_nonlocal_create, as the reference builds them
 * (PseudoPotOperator_mpi_v2.h:64-84, :108-166): l, E = energy_up, cutoff = cutoff_r[l], radial
 * kind QG_RADIAL_TABLE over the file's log grid with r_min = r[0] and the mean of the two
 * j-components.  specs[k].table points into v (valid until qg_vps_destroy).  cap >=
 * num_projectors. */
int qg_vps_projector_specs(const qg_vps* v, qg_proj_spec* specs, int cap);

typedef struct qg_nonlocal qg_nonlocal;

/* Build the projector geometry for natoms nuclei at xyz_host[3*a..3*a+2] (bohr), each of
 * species species_host[a]; species s owns nproj_host[s] consecutive entries of specs
 * (in species order).  Replaces PseudoPotIntegrator_mpi::UpdatePosition
 * (PseudoPotOperator_mpi_v2.h:231-248) for one process holding the whole grid:
 * GridForSphere boxes (GridSubgrid.h:483-492), their periodic overlap blocks
 * (GridSubgrid.h:381-430) and the projector/Ylm values on them. */
int qg_nonlocal_create(const qg_grid* g, int natoms, const double* xyz_host,
                       const int* species_host, int nspecies, const int* nproj_host,
                       const qg_proj_spec* specs, qg_nonlocal** out);
/* Move the nuclei of an existing handle: rebuild the projector geometry for natoms nuclei at
 * xyz_host (species indices into the channels given at creation) in place.  The handle keeps
 * its address, so every Hamiltonian built on it uses the new positions at its next
 * application.  Replaces PseudoPotIntegrator_mpi::UpdatePosition on a live integrator
 * (PseudoPotOperator_mpi_v2.h:231-248, called from qumasun_mpi.move.h:15 at every MD or
 * relaxation step).  Waits for the device first (no application may be in flight). */
int qg_nonlocal_update(qg_nonlocal* nl, int natoms, const double* xyz_host, const int* species_host);
int qg_nonlocal_destroy(qg_nonlocal* nl);
/* Number of (atom, projector, m) coefficients per band (sum of TotalProjectorLM). */
int qg_nonlocal_num_coef(const qg_nonlocal* nl);
/* How H psi forms the projector values: 1 = the FP32 projector pass evaluates them in the
 * kernel from the image centres (every projector Gaussian with l <= 1; QG_NL_ANALYTIC=0 in
 * the environment at creation turns it off), 0 = per-point tables everywhere. FP64 launches
 * and the fused single-pass form read tables. Diagnostic; no counterpart in the reference,
 * which always tabulates (SetBlockYlm.h). */
int qg_nonlocal_projector_form(const qg_nonlocal* nl);

/* out_x/y/z[b] = coef * (d/dx, d/dy, d/dz) psi[b], periodic 8th-order central differences
 * (Gradient8th / Gradient8th_ddm, GridGradient8th.h:17-160); psi and the three outputs are
 * band-major device blocks of one dtype (ld, ld_out >= Ng). */
int qg_gradient8_apply(const qg_grid* g, int dtype, int nbands, const void* psi, int64_t ld,
                       void* out_x, void* out_y, void* out_z, int64_t ld_out, double coef, void* stream);
/* Nonlocal energy and forces of occupied states (mGetEnergyPseudoNonlocal and
 * mGetForcePseudoNonlocal, qumasun_mpi.energy.h:349-375, qumasun_mpi.force.h:226-258, over
 * EnergyPPnonlocal / ForceNonlocal, PseudoPotOperator_mpi_v2.h:573-644, :905-1013): for the bands
 * with occ_host[b] > 0, *energy = sum_b occ_b sum_{a,j} E_j c_baj^2 and, when forces_host is not
 * NULL, forces_host[3a + d] = sum_b sum_j E_j c_baj (-2 occ_b) dV sum beta_aj d_d psi_b with the
 * 8th-order gradient (set, not accumulated).  psi: nbands band-major device vectors.  The handle
 * keeps the gradient scratch of a run of bands (at most 4 GiB) between calls, so calls on one
 * handle must not overlap. */
int qg_nonlocal_energy_force(qg_nonlocal* nl, int dtype, int nbands, const void* psi, int64_t ld,
                             const double* occ_host, double* energy, double* forces_host, void* stream);

/* coef[b*ncoef + j] = dV * sum_i beta_j(i) psi_b(i)  (FP64 device array, ncoef per band).
 * Replaces mInnerNonlocalPsi_v2 (PseudoPotOperator_mpi_v2.h:336-463) for a band block. */
int qg_nonlocal_project(qg_nonlocal* nl, int dtype, int nbands, const void* psi, int64_t ld,
                        double* coef_dev, void* stream);
/* hpsi[b] (+)= sum_j E_j coef_bj beta_j — ProjectionPP (PseudoPotOperator_mpi_v2.h:465-570)
 * and its band-block form ProjectionPP_bundle (:789-902), band-major layout. */
int qg_nonlocal_apply(qg_nonlocal* nl, int dtype, int nbands, const void* psi, int64_t ld,
                      void* hpsi, int64_t ld_out, int accumulate, void* stream);

/* --------------------------------------------------------------- Hamiltonian ---- */
typedef struct qg_ham qg_ham;

/* H = V (local, V_host Nx*Ny*Nz values, copied) + V_NL (nl, may be NULL) - 1/2 Lap8.
 * The operator of QUMASUN_MPI::mHamiltonianMatrix_ddm (qumasun_mpi.matrix.h:21-43). */
int qg_hamiltonian_create(const qg_grid* g, const double* V_host, qg_nonlocal* nl,
                          qg_ham** out);
int qg_hamiltonian_destroy(qg_ham* h);
/* Replace the local potential (SCF update of ml_Vtot). */
int qg_hamiltonian_set_potential(qg_ham* h, const double* V_host);
/* The same from a device array (FP64, Ng values): the SCF step installs the new potential
 * without leaving the device. */
int qg_hamiltonian_set_potential_device(qg_ham* h, const double* V_dev, void* stream);

/* -------------------------------------------------------------- k points ---- */
/* The k-point (complex Bloch) Hamiltonian of a real-space Hamiltonian h (QUMASUN_KPOINT,
 * qumasun_kpoint.matrix.h:21-110): for the cell-periodic part u of psi = exp(i k.r) u,
 *   H_k u = -1/2 [L8 u - |k|^2 u + 2 i k . D8 u] + V u + sum_j E_j conj(P_j) (dV P_j^T u),
 * P_j = beta_j(x_u - R) exp(i k . x_u) on the projector's image blocks (x_u the image-frame grid
 * coordinate; mCreateProjectorRadialOnBlock, PseudoPotOperator_kpoint_v2.h:287-420).  k in
 * bohr^-1 (kpoint_i * 2 pi / (L_i * samples_i) in the reference).  The handle keeps the
 * projector phase tables of h's nonlocal operator and rebuilds them when its geometry changes;
 * h must outlive it.  The handle also keeps its scratch (projector coefficients, FFT plans and
 * spectrum, the solver's BLAS handle) between calls, so calls on one handle must not overlap. */
typedef struct qg_kpoint qg_kpoint;
int qg_kpoint_create(qg_ham* h, double kx, double ky, double kz, qg_kpoint** out);
int qg_kpoint_destroy(qg_kpoint* kp);
/* hpsi[b] = H_k psi[b]: mHamiltonianMatrix_ddm at a k point (LaplasianKpoint8th_ddm,
 * GridDifferenceKpoint8th.h:17-148; ProjectionPP, PseudoPotOperator_kpoint_v2.h:2581-2690 over
 * mInnerNonlocalPsi_v2, :836-930).  dtype QG_Z128 or QG_C64; psi and hpsi band-major
 * interleaved complex, ld in complex elements (the reference's SoAComplex holds re and im as
 * separate arrays per band). */
int qg_kpoint_apply(qg_kpoint* kp, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi,
                    int64_t ld_out, void* stream);
/* Lowest nbands eigenpairs of H_k by block LOBPCG in complex FP64 (Eigen_z_multi_mpi,
 * lobpcg_z_multi_mpi.h:988-1600; ZGEMM Grams), on an orthonormal [X W P] basis (SVQB, standard
 * Hermitian Rayleigh-Ritz with rocSOLVER zheevd).  X: nbands QG_Z128 bands (ld = Ng), the start
 * on entry and the eigenvectors on return (dV-orthonormal); eig_host: nbands values.  Stops when
 * the largest residual norm sqrt(dV sum |H x - theta x|^2) over the lowest nconv bands (nconv <= 0:
 * all) is <= tol, or after iter_max iterations.  precond_shift <= 0: no preconditioner (the
 * reference's iteration); > 0: the residuals are preconditioned by (T_k + precond_shift)^-1, the
 * exact inverse of the shifted k-point kinetic operator (complex FFTs, rocFFT; no counterpart in
 * the reference).  Device workspace: 9 blocks of nbands complex FP64 bands ([X W P], their H
 * images, three work blocks). */
int qg_kpoint_lobpcg(qg_kpoint* kp, int nbands, void* X, int64_t ld, double* eig_host, int iter_max,
                     double tol, int nconv, double precond_shift, int* iters_out, double* max_res_out,
                     void* stream);

/* Layout conversion between the reference's complex band blocks (per band ng real parts then
 * ng imaginary parts, FP64: the SoAComplex arrays of Eigen_z_multi_mpi's IntervalPointers,
 * lobpcg_z_multi_mpi.h:1004-1011) and the engine's interleaved complex bands (dtype QG_Z128 or
 * QG_C64).  Device pointers; asynchronous on stream. */
int qg_complex_from_planar(int64_t ng, int nbands, const double* planar, int dtype, void* z, void* stream);
int qg_complex_to_planar(int64_t ng, int nbands, const void* z, int dtype, double* planar, void* stream);

/* ------------------------------------------------------------- SCF step ---- */
/* The steps on either side of the eigensolve in the SCF loop (QUMASUN_MPI::Execute,
 * qumasun_mpi.scfloop.h:79-160), spin-unpolarised, every field FP64 on the device.
 *
 * Occupancies (host): occ_s = 2 / (exp(min((e_s - mu) / kbT, 80)) + 1) with mu from the
 * bisection of mSetOccupancySpinOff (SetOccupancy.h:7-18, qumasun_mpi.occupancy.h:61-108; the
 * occupancies of the last bisection step; *mu_out the final bracket midpoint). */
int qg_occupancy_fermi(int nbands, const double* eig_host, double num_electrons, double kbT,
                       double* occ_host, double* mu_out);
/* rho = sum over bands with occ > 0 of occ * psi^2 (mSetDensityOne, qumasun_mpi.density.h:10-32);
 * psi band-major (ld >= Ng), QG_F64 or QG_F32; rho overwritten. */
int qg_density(const qg_grid* g, int dtype, int nbands, const void* psi, int64_t ld,
               const double* occ_host, double* rho_dev, void* stream);
/* Linear mixing (mSetDensity, qumasun_mpi.density.h:46-64): rho <- prev + ratio (rho - prev),
 * *diff_abs = dV sum |rho_in - prev|, *total = dV sum rho_out.  rho_prev_dev NULL: no mixing
 * (only *total).  Either output pointer may be NULL. */
int qg_density_mix(const qg_grid* g, double* rho_dev, const double* rho_prev_dev, double ratio,
                   double* diff_abs, double* total, void* stream);
/* Hartree potential of the periodic cell by FFT (SetPotentialByPoissonFFT_keepRhok,
 * poisson_fft.h:131-240): V_H(k) = 4 pi rho(k) / |k|^2, V_H(k = 0) = 0 (rocFFT, FP64). */
int qg_hartree_potential(const qg_grid* g, const double* rho_dev, double* vhart_dev, void* stream);
/* LDA exchange and correlation potentials (Perdew-Zunger 1981) of rho + pcc (pcc may be NULL),
 * and the energy densities e_x, e_c when ex_dev / ec_dev are not NULL (Calc_XC_LDA, Vxc.h:46-75;
 * zeros where rho + pcc < 1e-8). */
int qg_xc_lda(const qg_grid* g, const double* rho_dev, const double* pcc_dev, double* vx_dev,
              double* vc_dev, double* ex_dev, double* ec_dev, void* stream);
/* The potential step of the loop (mSetPotential, qumasun_mpi.potential.h:23-170 without the
 * high-resolution grid): V_H of rho, V_x and V_c of rho + pcc, V_tot = V_ext + V_H + V_x + V_c
 * (vext_dev may be NULL) installed in h and written to vtot_dev when not NULL. */
int qg_scf_potential(qg_ham* h, const double* rho_dev, const double* pcc_dev, const double* vext_dev,
                     double* vhart_dev, double* vx_dev, double* vc_dev, double* vtot_dev, void* stream);
/* hpsi[b] = H psi[b] for b < nbands in one fused pass (V*psi + nonlocal - 1/2 Lap8 psi).
 * Composition order of mHamiltonianMatrix_ddm (zero, V*psi, +ProjectionPP, +Laplacian). */
int qg_hamiltonian_apply(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld,
                         void* hpsi, int64_t ld_out, void* stream);
/* Same result as qg_hamiltonian_apply with no device memory allocated or freed, for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  It runs the row kernels only (Nx a
 * multiple of 4 in FP32 or of 2 in FP64, 8 <= Nx <= 1024, Gaussian projectors with l <= 1)
 * and keeps the projector partials and coefficients in one fixed 64 MiB device array per GPU,
 * processing the bands in chunks that fit it; QG_ERR_ARG when one band does not fit or the
 * projector geometry is not prepared (qg_nonlocal_create/update prepare it).  Calls on one GPU
 * share that array: calls that overlap on one GPU (other streams, other threads) race. */
int qg_hamiltonian_apply_noalloc(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld,
                                 void* hpsi, int64_t ld_out, void* stream);
/* Kernel plan of a Hamiltonian's H psi (no counterpart in the reference).  The row kernels
 * (a tile-centric projection launch and one stencil + V + projector launch) serve grids with
 * Nx a multiple of 4 (FP32) or 2 (FP64), 8 <= Nx <= 1024, and Gaussian projectors with l <= 1;
 * the table kernels serve everything.  Zero fields are chosen automatically from the grid and
 * the band count.  The QG_ROWS_* environment variables override the shape fields (measurement
 * switches). */
#define QG_KERNELS_AUTO 0     /* row kernels where they serve, table kernels otherwise */
#define QG_KERNELS_ROWS 1     /* row kernels only (QG_ERR_ARG where they do not serve) */
#define QG_KERNELS_TABLES 2   /* table kernels */
typedef struct qg_hpsi_plan {
    int kernels;      /* QG_KERNELS_* */
    int rows_ty;      /* apply kernel: rows per y-tile */
    int rows_bands;   /* apply kernel: bands per workgroup (1, 2 or 4) */
    int rows_tz;      /* apply kernel: z-planes per slab */
    int rows_tzp;     /* projection: z-planes per slab (default 8; the projector geometry of other
                         values is built at the first application) */
    int rows_nbp;     /* projection: bands per workgroup (4 or 8) */
} qg_hpsi_plan;
void qg_default_hpsi_plan(qg_hpsi_plan* plan);
int qg_hamiltonian_set_plan(qg_ham* h, const qg_hpsi_plan* plan);

/* Sector timers of qg_hamiltonian_apply on this handle (hipEvents recorded on the apply's
 * stream), the counterpart of the reference's StopWatch sectors 13 (PP nonlocal) and
 * 11/12 (kinetic, V) (qumasun_mpi.matrix.h:30-41): t_project_ms is the total time of the
 * projector-dot launches, t_fused_ms that of the fused stencil + V + projector-apply launches
 * of the last apply (one of each per band chunk).  last_times waits for the recorded events. */
int qg_hamiltonian_set_timing(qg_ham* h, int enable);
int qg_hamiltonian_last_times(qg_ham* h, double* t_project_ms, double* t_fused_ms);
/* Number of band chunks (= projection launches = fused launches) of the last timed apply. */
int qg_hamiltonian_last_chunks(qg_ham* h, int* nchunks);

/* ------------------------------------------------------------- block LOBPCG ---- */
/* precision modes */
#define QG_PREC_F64 0     /* everything FP64 (the reference's arithmetic) */
#define QG_PREC_MIXED 1   /* FP32 H·psi, BF16-in/FP32-acc MFMA Gram, FP64 small solve */
#define QG_PREC_F32 2     /* FP32 H·psi and FP32 Gram, FP64 small solve */

/* preconditioners (the reference uses the raw residual, lobpcg_d_multi_mpi.h:982-990) */
#define QG_PRECOND_NONE 0
#define QG_PRECOND_KINETIC 1   /* (T + shift)^-1 applied exactly in Fourier space */
#define QG_PRECOND_LEARNED 2   /* K + K N^T N K with the learned network attached to the
                                  Hamiltonian (qg_hamiltonian_set_cnn); falls back to
                                  QG_PRECOND_KINETIC when the residual stalls */

/* Rayleigh-Ritz basis handling.  QG_BASIS_REFERENCE: the reference's [X R P] generalized
 * problem with Gram recurrences and the rollback exit (FP64 without a preconditioner only).
 * QG_BASIS_ORTHO: [X W P] kept orthonormal (W = K R orthogonalised against X, P against X
 * and W, each orthonormalised by SVQB with dependent directions dropped), standard symmetric
 * Rayleigh-Ritz problem; always used with a preconditioner or in the FP32 modes. */
#define QG_BASIS_REFERENCE 0
#define QG_BASIS_ORTHO 1

/* Orthonormalisation of the W and P blocks in the orthonormal basis: QG_ORTH_CHOLQR (default)
 * Cholesky-QR of the diagonally scaled Gram matrix, with SVQB where the Cholesky factor fails or
 * its pivots show near-dependent columns; QG_ORTH_SVQB SVQB always (an eigensolve per block). */
#define QG_ORTH_CHOLQR 0
#define QG_ORTH_SVQB 1

typedef struct qg_lobpcg_opts {
    int precision;           /* QG_PREC_* */
    int max_iter;            /* iter_max of Eigen_d_multi_mpi */
    double tol;              /* stop when max_k ||r_k|| (dVol-normalised) <= tol; 0: run max_iter */
    int refresh_period;      /* full Gram recompute every n iterations (S_MAT_REST_LIMIT = 10) */
    double rollback_limit;   /* |1-||x'||^2| limit (LIMIT_D_NORM_X_FOR_RESET_MATRIX = 1e-3) */
    int precond;             /* QG_PRECOND_* */
    double precond_shift;    /* shift of the kinetic preconditioner (Ha) */
    int verbose;
    double switch_tol;       /* FP32 modes: residual at which FP64 refinement sweeps take over
                                (also when the FP32 residual stalls at its rounding floor: no
                                30% decrease within 3 * max(stall_window, 3) iterations) */
    int stall_window;        /* QG_PRECOND_LEARNED: iterations without a new minimum of the
                                block's residual norm (sqrt of the sum of the band residuals
                                squared) before the kinetic preconditioner takes over */
    /* Optional foreign operator (the reference's OpeA, lobpcg_d_multi_mpi.h:966-969): when
     * set, ax = A x for nbands band-major device vectors (ld == Ng, dtype QG_F64 or QG_F32)
     * is computed by this callback on the given stream instead of by the Hamiltonian; the
     * Hamiltonian still supplies the grid (and the preconditioner's operator). */
    void (*op)(void* user, int dtype, int nbands, const void* x_dev, void* ax_dev, void* stream);
    void* op_user;
    int basis;               /* QG_BASIS_*: used for FP64 runs without a preconditioner */
    double fallback_shift;   /* QG_PRECOND_LEARNED: shift of the kinetic preconditioner that
                                takes over after a stall (0: precond_shift) */
    /* Optional sector clock, the reference's WATCH::Record(idx, count) (StopWatch.h:34): called
     * on the solving thread once the device work of a sector has completed, with the sector ids
     * the reference's solver and operator use (qumasun_mpi.print.h:78-86,
     * lobpcg_d_multi_mpi.h:978-1571): 10 others (also called when a sector starts, so the time
     * between sectors is "others"), 11 H psi (kinetic, local and nonlocal in one application;
     * count = bands), 15 small eigensolve (LAPACK), 16 block updates (GEMM2), 17 Gram
     * products (GEMM1).  NULL: no calls. */
    void (*watch)(void* user, int sector, int64_t count);
    void* watch_user;
    int orth;                /* QG_ORTH_* (orthonormal basis only) */
    int gram_chunks;         /* K chunks of the split-K library Gram products (each chunk summed
                                in the block precision, the chunks in FP64); 0: 64 in FP64,
                                chunks of at most ~6,750 terms in FP32 (256 at 120^3: longer FP32
                                partial sums stalled the FP32 phase at 512 bands) */
    double reortho;          /* orthonormal basis: a second orthonormalisation pass runs when the
                                first one's pivot (or SVQB) ratio is at or below this; 0: 1e-3 in
                                FP64, always in FP32 */
    int nconv;               /* the residual tests (tol, switch_tol, stall_window, the reported
                                max_residual) look at the lowest nconv bands only; the others are
                                guard bands that keep the block's edge off a degenerate multiplet
                                of the wanted ones.  0: all bands (the reference's behaviour) */
} qg_lobpcg_opts;

typedef struct qg_lobpcg_stats {
    int iterations;
    int rollbacks;
    int hpsi_calls;          /* band applications of H */
    double max_residual;     /* final max_k ||r_k|| over the lowest nconv bands (dVol-normalised) */
    double t_hpsi, t_gram, t_update, t_eig, t_other, t_precond;   /* seconds per sector */
    int refine_iterations;   /* FP64 refinement iterations of an FP32-mode run */
    int fallback_iter;       /* iteration at which a stalled learned preconditioner was
                                replaced by the kinetic one (-1: never) */
} qg_lobpcg_stats;

void qg_lobpcg_default_opts(qg_lobpcg_opts* o);

/* Block LOBPCG for the nbands lowest eigenpairs of H; the device-resident equivalent of
 * LOBPCG::Eigen_d_multi_mpi (lobpcg_d_multi_mpi.h:882-1572) on one process.
 *   X_dev: nbands x Ng FP64 band-major, in: start vectors (normalised with dVol), out:
 *          Ritz vectors normalised with dVol (the reference's eigen_vectors).
 *   P_dev: nbands x Ng FP64, the persistent conjugate block (keep_P); zeroed when
 *          scf_step == 0 (lobpcg_d_multi_mpi.h:992-997).
 *   eigvals_host: nbands Ritz values, ascending.
 * Work memory is owned by the handle (the reference's `work`, 5*Ng*N doubles). */
int qg_lobpcg_d_multi(qg_ham* h, const qg_lobpcg_opts* opts, int nbands,
                      double* eigvals_host, double* X_dev, double* P_dev, int scf_step,
                      qg_lobpcg_stats* stats);

/* Orthonormalise nbands band-major device vectors X (ld >= ng, FP64 or FP32) in place so
 * that dVol * sum_i x_m(i) x_n(i) = delta_mn, with the result of Gram-Schmidt in band order.
 * Replaces GramSchmidt (GramSchmidt.h:6-38) and GramSchmidt_ddm (GramSchmidt_mpi.h:9-53),
 * used on the initial states (qumasun_mpi.initialize.h:183, :253): computed as two
 * Cholesky-QR passes (Gram GEMM, Cholesky, triangular solve), which give the same basis as
 * the band-by-band projections up to rounding.  QG_ERR_NUMERIC when the bands are dependent. */
int qg_gram_schmidt(int dtype, int64_t ng, int nbands, void* X, int64_t ld, double dvol, void* stream);

/* Apply a preconditioner in place to nbands residual vectors r (device, ld == Ng):
 * QG_PRECOND_KINETIC: r <- (T + shift)^-1 r with T = -1/2 Lap8 the periodic 8th-order kinetic
 * operator, exactly, through FFTs.  No counterpart in the reference (which uses the raw
 * residual, lobpcg_d_multi_mpi.h:982-990); it is the classic preconditioner stage. */
int qg_precond_apply(qg_ham* h, int kind, double shift, int dtype, int nbands, void* r, int64_t ld,
                     void* stream);

/* Gram product C = A^T B of two tall band blocks in their own precision (dtype QG_F64 or
 * QG_F32): A (K x M, column m at A + m*lda) and B (K x N) device blocks, C FP64 device (column
 * stride ldc).  Split-K library GEMMs in the block precision with the chunks summed in FP64:
 * the arithmetic of the solver's Rayleigh-Ritz Gram blocks, which replace blas_DGEMM_t
 * (w_dgemm.h:32-37) inside MakeMatrix_d_blas (lobpcg_d_multi_mpi.h:192-250).  Allocates its
 * workspace and waits for the stream. */
int qg_gram(int dtype, const void* A, const void* B, int64_t K, int M, int N, int64_t lda, int64_t ldb,
            double* C, int ldc, void* stream);

/* Gram product C = A^T B on the matrix cores: A (K x M) and B (K x N) FP32 device blocks
 * (column m at A + m*lda), rounded to BF16 (split != 0: x = hi + lo in two BF16 terms,
 * three MFMAs per step), FP32 accumulation, C FP64 device (column stride ldc).  The
 * Rayleigh-Ritz product of the mixed-precision LOBPCG; replaces blas_DGEMM_t
 * (w_dgemm.h:32-37) inside MakeMatrix_d_blas (lobpcg_d_multi_mpi.h:192-250). */
int qg_gram_bf16(const float* A, const float* B, int64_t K, int M, int N, int64_t lda, int64_t ldb,
                 double* C, int ldc, int split, void* stream);

/* ------------------------------------------------- learned preconditioner ---- */
/* A 3-layer 3x3x3 convolution network N on the periodic grid (8 hidden channels, inputs: the
 * residual and the local-potential patch, V entering through tanh gates), used as
 *   M r = K r + K N^T N K r,   K = (T + shift)^-1,
 * which is symmetric positive definite for any weights.  Trained on the device by this
 * library's forward/backward kernels on pairs r -> (T + V - sigma)^-1 r that the solver's own
 * residual blocks and a preconditioned CG produce.  No counterpart in the reference (which
 * has no preconditioner, lobpcg_d_multi_mpi.h:982-990). */
typedef struct qg_cnn qg_cnn;
int qg_cnn_num_params(void);
/* Network for Hamiltonian h (its V is the gate input), kinetic shift s, room for max_pairs
 * training pairs, weights drawn from a seeded normal distribution. */
int qg_cnn_create(qg_ham* h, double shift, int max_pairs, uint64_t seed, qg_cnn** out);
int qg_cnn_destroy(qg_cnn* c);
int qg_cnn_get_params(qg_cnn* c, float* host, int n);
int qg_cnn_set_params(qg_cnn* c, const float* host, int n);
/* out = N in (adjoint == 0) or N^T in (adjoint != 0) for nbands FP32 device vectors of Ng;
 * bf16 != 0 uses the BF16-rounded weights (the preconditioner's), else the FP32 master copy. */
int qg_cnn_network(qg_cnn* c, int adjoint, int bf16, int nbands, const float* in, float* out,
                   void* stream);
/* Loss and gradient for one batch: u, t are nbands x Ng FP32 device vectors, znorm the
 * normaliser; loss = sum |K N^T N u - t|^2 / znorm, gradient (qg_cnn_num_params floats) to
 * grad_host. */
int qg_cnn_loss_grad(qg_cnn* c, int nbands, const float* u, const float* t, double znorm,
                     double* loss, float* grad_host);
/* Add training pairs from solver iterates X (nbands x Ng FP64 device) with Ritz values
 * eig_host: r = H x - e x, z = (T + V - sigma)^-1 r by PCG to relative tolerance tol. */
int qg_cnn_collect(qg_cnn* c, int nbands, const double* X_dev, const double* eig_host,
                   double sigma, double tol, int maxit, int* cg_iters);
int qg_cnn_num_pairs(const qg_cnn* c);
/* Adam on random batches of the stored pairs; mean loss of the first and last 10% steps. */
int qg_cnn_train(qg_cnn* c, int steps, int batch, double lr, uint64_t seed,
                 double* loss_first, double* loss_last);
/* r <- M r in place (nbands device vectors, ld == Ng, FP32 or FP64). */
int qg_cnn_apply(qg_cnn* c, int dtype, int nbands, void* r, int64_t ld, void* stream);
/* Attach (or detach with NULL) the network used by QG_PRECOND_LEARNED. */
int qg_hamiltonian_set_cnn(qg_ham* h, qg_cnn* c);

/* ------------------------------------------------- batches of uneven cells ---- */
/* H psi for ncells independent cells (each its own Hamiltonian, grid, atoms and band count;
 * psi[c], hpsi[c] band-major device blocks with ld == Ng of that cell) in persistent launches:
 * one projection launch and one fused stencil/V/projector launch per round of Infinity-Cache
 * sized band chunks, covering every cell.  Same results as qg_hamiltonian_apply per cell.
 * The cells must share the projector slot count (cells with different tile plans run in
 * one stencil launch per plan).  No
 * counterpart in the reference (which runs one cell per process). */
int qg_batch_hamiltonian_apply(int ncells, qg_ham* const* hams, int dtype, const int* nbands,
                               const void* const* psi, void* const* hpsi, void* stream);
/* Autotuner: times the fused H psi of h on nbands zero bands for candidate tile shapes (rows
 * per tile x points per thread; fixed_ppt != 0 pins the latter) and records the fastest for
 * this grid shape and dtype; later applications on grids of that shape use it (unless the
 * QG_PLAN_TY / QG_PLAN_PPT environment overrides are set). */
int qg_autotune_hamiltonian(qg_ham* h, int dtype, int nbands, int fixed_ppt, int* ty, int* ppt,
                            double* best_ms);
/* LOBPCG for ncells independent cells (arguments per cell as in qg_lobpcg_d_multi).  batched
 * != 0: the cells are solved together, each on its own host thread and stream, and the H psi
 * requests of all cells meet and run as one qg_batch_hamiltonian_apply per round; batched ==
 * 0: one cell after another.  stats (may be NULL) receives ncells records. */
int qg_batch_lobpcg(int ncells, qg_ham* const* hams, const qg_lobpcg_opts* opts, const int* nbands,
                    double* const* eigvals_host, double* const* X_dev, double* const* P_dev,
                    int scf_step, qg_lobpcg_stats* stats, int batched);

/* ------------------------------------------------------- synthetic workloads ---- */
/* Host-only helpers that build the seeded diamond cells used to exercise this path:
 * ncell^3 diamond supercell of the 8-atom cubic cell with lattice a (bohr) on an
 * (n*ncell)^3 grid.  Atom fractional sites follow input/diamond8.txt:28-35. */
int qg_synth_diamond_atoms(int ncell, double a, double* xyz_host /* 24*ncell^3 */);
/* V(r) = sum_atoms -zion*erf(|r-R|_min/rc)/|r-R|_min + U(-noise, noise), minimum image,
 * noise from std::mt19937(seed). */
int qg_synth_vloc(int nx, int ny, int nz, double lx, double ly, double lz, int natoms,
                  const double* xyz_host, double zion, double rc, double noise,
                  uint32_t seed, double* V_host);
/* Random start block: U[0,1) from std::mt19937(seed), each band normalised with dVol
 * (qumasun_mpi.initialize.h:199-214). */
int qg_synth_random_states(int64_t ng, int nbands, double dvol, uint32_t seed,
                           double* X_host);

#ifdef __cplusplus
}
#endif
#endif
