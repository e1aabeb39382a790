// qg_internal.hpp — shared host/device declarations of the MI355X engine (not part of the
// public C ABI in include/qumasun_gpu.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include <cstdlib>
#include <memory>
#include <mutex>
#include "../../include/qumasun_gpu.h"

namespace qg {

// ---------------------------------------------------------------- error plumbing ----
void set_error(const std::string& msg);
struct Error {
    int code;
    std::string msg;
};
#define QG_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::qg::Error{QG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)
#define QG_REQUIRE(cond, msg)                                         \
    do {                                                              \
        if (!(cond)) throw ::qg::Error{QG_ERR_ARG, std::string(msg)}; \
    } while (0)

// Run f, converting exceptions into a status code + qg_last_error().
template <class F>
int guarded(F&& f) {
    try {
        f();
        return QG_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return QG_ERR_ARG;
    }
}

hipStream_t default_stream();
// NULL is HIP's null stream (the legacy default stream), as everywhere in HIP.
inline hipStream_t as_stream(void* s) { return (hipStream_t)s; }

// ------------------------------------------------------------- device buffers ----
// Integer tuning override from the environment (dflt when unset or empty).
inline int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return (v && *v) ? std::atoi(v) : dflt;
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count <= n && p) return;
        release();
        if (count == 0) return;
        QG_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T* host, size_t count, hipStream_t s = nullptr) {
        alloc(count);
        if (count) QG_HIP(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
        if (count) QG_HIP(hipStreamSynchronize(s));
    }
};

}  // namespace qg

// ------------------------------------------------------------------ handles ----
struct qg_grid {
    int n[3];
    double h[3];
    int64_t ng() const { return (int64_t)n[0] * n[1] * n[2]; }
    double dvol() const { return h[0] * h[1] * h[2]; }
};

namespace qg {

// One projector channel as the kernels see it.
struct DevProj {
    int l;
    int kind;          // QG_RADIAL_*
    int coef_rel;      // coefficient index of this channel's m = first within its atom
    int spec_index;    // index into the caller's qg_proj_spec array
    int table_off;     // offset into the table pool (QG_RADIAL_TABLE)
    int n_table;
    double energy;
    double amp;        // gaussian amplitude
    double inv_2s2;    // 1/(2 sigma^2)
    double xi_min, xi_delta, r_min2;
};

// One periodic-image block of one (atom, cutoff group): the overlap of the atom's
// GridForSphere box shifted by (sx,sy,sz) grid periods with the grid.
struct DevImage {
    int b[3], e[3];     // in-grid range [b, e)
    int shift[3];       // grid-index shift of the image (multiples of N)
    int row_base;       // offset of this block's first row in the interval table
    int coef_base;      // first coefficient of the group (atom-major coefficient order)
    int proj_begin;     // first DevProj of the group
    int proj_end;
    int atom;
    double R[3];        // nucleus position
};

// Host-side geometry of the nonlocal operator (built once per UpdatePosition).
struct NlGeometry {
    std::vector<DevProj> projs;
    std::vector<double> tables;
    std::vector<DevImage> images;
    std::vector<int2> intervals;   // per row of each image block: [x0, x1) inside the cutoff
    std::vector<int> atom_img_off; // images of atom a: [atom_img_off[a], atom_img_off[a+1])
    std::vector<int> atom_coef_off;
    std::vector<double> coef_energy;   // per coefficient: the KB energy of its projector
    int ncoef = 0;
    int natoms = 0;
    int max_coef_per_atom = 0;
    // projector tables on the sphere points (the reference's per-block projector x Ylm
    // tables), flattened image by image in (z, y, x) order
    int njs = 4;                       // table stride: coefficient slots per point (4 or 8)
    std::vector<int64_t> img_tbl_off;  // points of image i: [img_tbl_off[i], img_tbl_off[i+1])
    std::vector<int64_t> row_tbl_off;  // per interval row: table index of its first point
    std::vector<double> beta;          // npts * njs, slot j = coefficient within the atom
    std::vector<int> gidx;             // grid index of every table point
};

void build_nl_geometry(const qg_grid& g, int natoms, const double* xyz, const int* species,
                       int nspecies, const int* nproj, const qg_proj_spec* specs, NlGeometry& out);

// Projector data the fused kernel stages per workgroup: for every (y tile, z slab) the
// images and atoms that reach it, and for every (y tile, z plane) either the table items
// (point, atom slot, projector values) or the sphere segments of each tile row.
struct TileNL {
    int ty = 0, tz = 0, ntile = 0, nslab = 0;
    std::vector<int> wg_img_off, wg_img;     // per (tile, slab): global image ids
    std::vector<int> wg_img_atom;            // parallel to wg_img: local atom slot
    std::vector<int> wg_atom_off, wg_atom;   // per (tile, slab): global atom ids
    // work items of the projector apply: one per (image, sphere point), grouped by
    // (tile, plane), each with its tile-local point index and the workgroup's atom slot
    std::vector<int64_t> item_off;           // per (tile, plane): [item_off[tp], item_off[tp+1])
    std::vector<uint32_t> item_meta;         // tile point | atom slot << 16
    std::vector<double> item_beta;           // njs table values per item
    // sphere segments of every tile row: per (tile, plane, row) [seg_off[k], seg_off[k+1]),
    // each x0 | x1 << 11 | local image << 22 (the in-cutoff x range of one image on the row)
    std::vector<int> seg_off;
    std::vector<uint32_t> seg;
    int max_img = 0, max_atom = 0, max_items = 0, max_seg = 0;   // max_seg: per (tile, slab)
};
void build_tile_nl(const qg_grid& g, const NlGeometry& geo, int ty, int tz, bool segments, TileNL& out);

// Geometry of the row kernels (qg_rows.hip) for y-tiles of ty rows and projection slabs of
// tzp planes: the periodic sphere images of every tile column and the sphere points of every
// (tile, plane) as items (row | x start << 5 | (count - 1) << 15 | column image << 17).
// Projection order: items of up to 4 consecutive points grouped by image, with the
// projection's per-thread image ownership and its partial slots.  Group order (the apply
// kernel): items cut at multiples of 4 in x, so each lies in one 4-point group of its row,
// sorted by (row, group, image), with the offsets of every group's run of items.
struct RowsGeo {
    int ty = 0, tzp = 0, ntile = 0, nslabp = 0;
    bool by_group = false;
    int ng4 = 0;         // 4-point groups per row (group order)
    int max_col = 0;     // most images of one tile column
    int max_items = 0;   // most items of one (tile, plane)
    int nslot = 0;       // projection partial slots (one per active (tile, slab, image))
    int njs = 4;         // coefficient slots per atom (4 or 8)
    bool sp4 = false;    // every image: one s channel (slot 0) and one p channel (slots 1..3)
    int pform = 0;       // projector form of the kernels: 0 general, 1 s + p, 2 s + p with one exponent
    bool ok = false;     // the geometry fits the encodings
    std::vector<int> col_off, col_img, it_off, pl_base, pl_off, thr_slot, atom_slot_off, atom_slot;
    std::vector<int> grp_off;   // group order: per (tile, plane) ty * ng4 + 1 run offsets (plane-relative)
    std::vector<uint32_t> items, thr_map;
    DevBuf<int> d_col_off, d_col_img, d_it_off, d_pl_base, d_pl_off, d_thr_slot, d_atom_slot_off, d_atom_slot;
    DevBuf<int> d_grp_off;
    DevBuf<uint32_t> d_items, d_thr_map;
};

// Threads of a projection workgroup of the row kernels (the thread map is built for it).
constexpr int kRowsProjThreads = 256;
bool build_rows_geometry(const qg_grid& g, const NlGeometry& geo, int ty, int tzp, bool by_group, RowsGeo& o);

// Caller-provided FP64 scratch of one application (projector partials, then coefficients).
struct RowsScratch {
    double* base;
    size_t capacity;   // doubles
};

}  // namespace qg

struct qg_nonlocal {
    const qg_grid* grid = nullptr;
    qg_grid gcopy;
    // the projector channels given at creation (tables copied), reused by qg_nonlocal_update
    int nspecies = 0;
    std::vector<int> nproj;
    std::vector<qg_proj_spec> specs;
    std::vector<std::vector<double>> spec_tables;
    qg::NlGeometry geo;
    qg::DevBuf<qg::DevProj> d_projs;
    qg::DevBuf<double> d_tables;
    qg::DevBuf<qg::DevImage> d_images;
    qg::DevBuf<int2> d_intervals;
    qg::DevBuf<int> d_atom_img_off;
    qg::DevBuf<int> d_atom_coef_off;
    qg::DevBuf<double> d_coef;   // scratch coefficients (nbands * ncoef)
    // per-plan staging lists, cached for one (ty, tz) per projector form: [0] per-point
    // tables (item lists), [1] in-kernel values (row segments)
    struct Tiles {
        int ty = -1, tz = -1;
        qg::TileNL tile;
        qg::DevBuf<int> d_wg_img_off, d_wg_img, d_wg_img_atom, d_wg_atom_off, d_wg_atom;
        qg::DevBuf<int64_t> d_item_off;
        qg::DevBuf<uint32_t> d_item_meta;
        qg::DevBuf<int> d_seg_off;
        qg::DevBuf<uint32_t> d_seg;
        qg::DevBuf<double> d_item_beta64;
        qg::DevBuf<float> d_item_beta32;
    } tiles[2];
    // every projector Gaussian with l <= 1 (and QG_NL_ANALYTIC != 0 at creation): the FP32
    // projector pass evaluates the values in the kernel
    bool analytic = false;
    int analytic_level = 0;   // QG_NL_ANALYTIC: 1 FP32 projector pass, 2 FP64 projector pass too
    qg::DevBuf<double> d_coef_energy;   // KB energy of every coefficient slot
    // projector tables
    qg::DevBuf<double> d_beta64;
    qg::DevBuf<float> d_beta32;
    qg::DevBuf<int> d_gidx;
    qg::DevBuf<int64_t> d_img_tbl_off;
    // row-kernel geometry, one per (tile rows, projection slab) shape in use
    std::vector<std::unique_ptr<qg::RowsGeo>> rows;
    std::mutex rows_mu;             // guards the creation of entries of `rows`
    qg::DevBuf<double> rows_part;   // projector partials of qg_nonlocal_project
    // geometry generation: bumped by every (re)build of the projector geometry, so state derived
    // from it elsewhere (the k-point phase tables) knows when to rebuild
    int geo_version = 0;
    // qg_nonlocal_energy_force: gradient blocks of a band run and their coefficients
    qg::DevBuf<unsigned char> force_grad;
    qg::DevBuf<double> force_coef;
};

namespace qg {
struct LobpcgState;
struct PrecondPlan;
struct BatchService;
void lobpcg_state_free(LobpcgState*);
void precond_state_free(PrecondPlan*);
bool rows_hamiltonian_apply(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi,
                            int64_t ld_out, hipStream_t s, RowsScratch* scratch);
bool rows_nonlocal_project(qg_nonlocal* nl, int dtype, int nbands, const void* psi, int64_t ld, double* coef,
                           DevBuf<double>& part_buf, hipStream_t s);
bool rows_laplacian_apply(const qg_grid& g, int dtype, int nbands, const void* psi, int64_t ld, void* out,
                          int64_t ld_out, double coef, int accumulate, hipStream_t s);
void rows_prepare(qg_nonlocal* nl);
double* scratch_address();
void rows_apply_noalloc(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi, int64_t ld_out,
                        hipStream_t s);
}  // namespace qg

struct qg_ham {
    qg_grid grid;
    qg_hpsi_plan plan{};     // kernel plan of H psi (zeros: automatic)
    qg_cnn* cnn = nullptr;   // learned preconditioner (not owned)
    qg::BatchService* batch = nullptr;   // set while a batched solve runs this cell
    qg::LobpcgState* lobpcg = nullptr;
    qg::PrecondPlan* precond = nullptr;
    qg_nonlocal* nl = nullptr;
    qg::DevBuf<double> V64;
    qg::DevBuf<float> V32;
    bool timing = false;
    bool timed_valid = false;
    std::vector<hipEvent_t> tev;   // per band chunk: before projection, before fused, after fused
    qg::DevBuf<double> rows_part;  // projector partials of the row path
    qg::DevBuf<double> rows_coef;  // projector coefficients of the row path
    int tchunks = 0;
    ~qg_ham() {
        for (auto& e : tev)
            if (e) (void)hipEventDestroy(e);
        if (lobpcg) qg::lobpcg_state_free(lobpcg);
        if (precond) qg::precond_state_free(precond);
    }
};
