// qg_rows.hip — H·psi on MI355X (gfx950) with 16-byte row vectors: a tile-centric projector
// dot-product kernel and a stencil + local potential + projector-apply kernel that writes each
// H psi value exactly once.
//
// Reference semantics (QUMASUN source/):
//   out += coef Lap8 psi, periodic            GridDifference8th.h:83-185
//   H psi = V psi + V_NL psi - 1/2 Lap8 psi   qumasun_mpi.matrix.h:21-43
//   c_aj = dV sum_i beta_j(i) psi(i)          PseudoPotOperator_mpi_v2.h:336-463
//   V_NL psi = sum_aj E_j c_aj beta_j         PseudoPotOperator_mpi_v2.h:465-570
//   beta_j = p(r) Y_lm(r), Y_1m = sqrt(3/4pi) (y, z, x)/r, Y_1m(r^2 < 1e-14) = 0
//                                             RealSphericalHarmonics.h:149-161, SetBlockYlm.h:41
//
// Geometry.  A y-tile is TY full x-rows of every z-plane.  Its "column" images are the
// periodic sphere images whose y-range meets the tile.  The sphere points of one (tile,
// plane) are listed as items: up to 4 consecutive points of one image on one tile row
// (row | x start << 5 | (count - 1) << 15 | column image << 17), sorted by image.
//
// Projection (rows_project_kernel): a workgroup owns a tile, a slab of planes and NBP bands.
// Per plane it stages the tile's rows of its bands in LDS (band fastest, one HBM read of
// psi); every thread owns one column image for the whole slab (threads per image in
// proportion to its sphere points) and accumulates beta . psi in registers.  At the slab's
// end the image's threads reduce through LDS and write one FP64 partial per (slot, band);
// rows_coef_reduce_kernel sums an atom's partials (no global atomics, deterministic).
//
// Apply (rows_hpsi_kernel): a workgroup owns a tile, a slab of planes and NBW bands (one
// group of threads per band); each thread owns VN = 16 / sizeof(T) consecutive x-points of
// one row and streams through the slab in z with the 9 z-neighbours in registers, the
// current plane (+ 4 wrapped columns each side and the 4-row y-halo) in LDS, double
// buffered.  Per plane the workgroup first evaluates the projector term item-parallel
// (beta computed once per sphere point and applied to all NBW bands, LDS float adds into a
// per-plane buffer), then each thread forms -1/2 Lap8 psi + V psi + that term and stores it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include "qg_internal.hpp"

namespace qg {

namespace {

using rsrc_t = __amdgpu_buffer_rsrc_t;
using u4 = __attribute__((ext_vector_type(4))) uint32_t;

template <class T>
struct alignas(16) Row {
    T v[16 / sizeof(T)];
};

__device__ __forceinline__ rsrc_t rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
template <class T>
__device__ __forceinline__ Row<T> vload(rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(Row<T>, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
template <class T>
__device__ __forceinline__ void vstore(rsrc_t r, const Row<T>& v, int voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, voff, soff, 0);
}

__device__ __forceinline__ int xcd_order(int i, int n) {
    // hardware workgroup ids go round-robin over the 8 XCDs; give each XCD one contiguous
    // range of logical ids (neighbouring tiles of a band share halo rows in that XCD's L2)
    const int xcd = i & 7, slot = i >> 3;
    const int per = n >> 3, rem = n & 7;
    return ((xcd < rem) ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + slot;
}

template <class T>
struct Coef8 {
    T c0;
    T cx[4], cy[4], cz[4];
};

// exp(-rr k') with k' = k log2(e) in FP32 (the hardware exp2, v_exp_f32: results below the
// normal range flush to zero, ~1e-38 of the amplitude), exp(-rr k) in FP64
template <class T>
__device__ __forceinline__ T gauss(T rr, T k) {
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_exp2f(-(rr * k));
    else return exp(-(rr * k));
}
// 1/sqrt(rr): the hardware FP32 estimate; in FP64 that estimate refined by two Newton steps
// y <- y (3 - rr y^2) / 2 (relative error ~1e-7 -> ~1e-14 -> rounding), rr in [1e-14, 1e30]
template <class T>
__device__ __forceinline__ T inv_sqrt(T rr) {
    if constexpr (sizeof(T) == 4) {
        return __frsqrt_rn(rr);
    } else {
        double y = (double)__frsqrt_rn((float)rr);
        y = y * (1.5 - 0.5 * rr * y * y);
        y = y * (1.5 - 0.5 * rr * y * y);
        return y;
    }
}

}  // namespace

// Per-image staging of the projector values: integer part n of the centre in grid units
// (shift included) and fractional part f, so the displacement h((i - n) - f) is exact at a
// nucleus on a grid point; projector range; for the s+p form the scaled amplitudes.
struct RowsImg {
    int n[3];
    int pb, pe;
    int atom;
};

struct RowsHpsiArgs {
    int nx, ny, nz;
    int ty, tz, ntile, nslab, nbands;
    int tpb;                    // threads per band (whole wavefronts)
    int64_t ld_in, ld_out;
    int add_v;                  // H psi: + V psi
    int accumulate;             // Laplacian only: + previous out
    int has_nl;
    int njs;                    // coefficient slots per atom in the staged E*c (4 or 8)
    const int* col_off;         // per tile: first column image
    const int* col_img;         // global image ids
    const int* it_off;          // per (tile, plane): first item
    const uint32_t* items;
    int nitems;                 // total items (item loads are clamped to the array)
    const int* grp_off;         // per (tile, plane): ty * ng4 + 1 run offsets of the 4-point groups
    int ng4;                    // 4-point groups per row
    const DevImage* images;
    const DevProj* projs;
    int nproj;
    const double* coef;         // nbands x ncoef, this application's bands
    const double* coef_energy;
    const int* atom_coef_off;
    int ncoef;
    int max_col;
    double h[3];
};

struct RowsProjArgs {
    int nx, ny, nz;
    int ty, tz, ntile, nslab, nbands;
    int64_t ld;
    const int* col_off;
    const int* col_img;
    const int* it_off;
    const uint32_t* items;
    int nitems;                 // total items (item loads are clamped to the array)
    const int* pl_base;         // per tile: base of its (plane, column image) item offsets
    const int* pl_off;          // (plane, image) -> first item of the image in the plane
    const uint32_t* thr_map;    // per (tile, slab, thread): image | rank << 10 | (count-1) << 18
    const int* thr_slot;        // per (tile, slab, thread): partial slot of the thread's image
    const DevImage* images;
    const DevProj* projs;
    int nproj;
    double* part;               // part[(slot * NJS + j) * part_ld + band]
    int part_ld;
    int max_col;
    double h[3];
};

// Projector-value staging shared by both kernels: image integer/fraction parts, and per
// projector l, coefficient slot, amplitude with its Y_lm constant, and exponent factor.
template <class T>
struct ProjLds {
    T* im_q;       // [max_col][8]: f.x f.y f.z, s amplitude * Y00, s exponent, p amplitude * sqrt(3/4pi),
                   // p exponent, 0 (the s + p forms; 16-byte rows)
    int* im_n;     // [max_col][4]: n.x n.y n.z pb (16-byte rows)
    int* im_i;     // [max_col][5]: n.x n.y n.z pb pe
    T* im_f;       // [max_col][3]: f
    int* pr_i;     // [nproj][2]: l, coef_rel
    T* pr_f;       // [nproj][2]: amp * (Y00 or sqrt(3/4pi)), exponent factor
};

// LDS bytes of the projector staging (ProjLds) for max_col images and nproj projectors; the
// general arrays (im_i, im_f, pr_*) only with gen (the general projector form)
template <class T>
__host__ __device__ constexpr size_t proj_lds_bytes(size_t max_col, size_t nproj, bool gen) {
    return (8 * sizeof(T) + 16) * max_col +
           (gen ? sizeof(T) * (3 * max_col + 2 * nproj) + sizeof(int) * (5 * max_col + 2 * nproj) : 0);
}

// carve the projector staging out of LDS at lp (16-byte aligned); GEN: with the general arrays
template <class T, bool GEN>
__device__ __forceinline__ ProjLds<T> carve_proj(unsigned char* lp, int max_col, int nproj) {
    ProjLds<T> L;
    L.im_q = reinterpret_cast<T*>(lp);
    lp += sizeof(T) * 8 * max_col;
    L.im_n = reinterpret_cast<int*>(lp);
    lp += sizeof(int) * 4 * max_col;
    if constexpr (!GEN) {
        L.im_f = nullptr;
        L.pr_f = nullptr;
        L.im_i = nullptr;
        L.pr_i = nullptr;
        return L;
    }
    L.im_f = reinterpret_cast<T*>(lp);
    lp += sizeof(T) * 3 * max_col;
    L.pr_f = reinterpret_cast<T*>(lp);
    lp += sizeof(T) * 2 * nproj;
    L.im_i = reinterpret_cast<int*>(lp);
    lp += sizeof(int) * 5 * max_col;
    L.pr_i = reinterpret_cast<int*>(lp);
    return L;
}

// amplitude with its Y_lm constant, and exponent factor of projector d (log2(e) folded in for
// the FP32 exp2)
template <class T>
__device__ __forceinline__ void proj_consts(const DevProj& d, T& amp, T& k) {
    const double ylm = d.l == 0 ? 0.28209479177387814 : 0.48860251190291992;   // 1/sqrt(4pi), sqrt(3/4pi)
    amp = (T)(d.amp * ylm);
    k = (sizeof(T) == 4) ? (T)(d.inv_2s2 * 1.4426950408889634) : (T)d.inv_2s2;
}

template <class T, bool GEN>
__device__ __forceinline__ void stage_proj(const ProjLds<T>& L, const DevImage* images, const DevProj* projs,
                                           int nproj, const int* col_img, int ncol, const double* h, int tid,
                                           int nthr) {
    for (int li = tid; li < ncol; li += nthr) {
        const DevImage& im = images[col_img[li]];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double u = im.R[d] / h[d];
            const double nn = rint(u);
            L.im_q[8 * li + d] = (T)(u - nn);
            L.im_n[4 * li + d] = im.shift[d] + (int)nn;
            if constexpr (GEN) {
                L.im_f[3 * li + d] = (T)(u - nn);
                L.im_i[5 * li + d] = im.shift[d] + (int)nn;
            }
        }
        if constexpr (GEN) {
            L.im_i[5 * li + 3] = im.proj_begin;
            L.im_i[5 * li + 4] = im.proj_end;
        }
        L.im_n[4 * li + 3] = im.proj_begin;
        T as = T(0), ks = T(0), ap = T(0), kp = T(0);
        if (im.proj_end > im.proj_begin) proj_consts<T>(projs[im.proj_begin], as, ks);
        if (im.proj_end > im.proj_begin + 1) proj_consts<T>(projs[im.proj_begin + 1], ap, kp);
        L.im_q[8 * li + 3] = as;
        L.im_q[8 * li + 4] = ks;
        L.im_q[8 * li + 5] = ap;
        L.im_q[8 * li + 6] = kp;
        L.im_q[8 * li + 7] = T(0);
    }
    if constexpr (GEN) {
        for (int p = tid; p < nproj; p += nthr) {
            const DevProj& d = projs[p];
            L.pr_i[2 * p] = d.l;
            L.pr_i[2 * p + 1] = d.coef_rel;
            proj_consts<T>(d, L.pr_f[2 * p], L.pr_f[2 * p + 1]);
        }
    }
}

// s + p projector values at 4 consecutive points (x offsets u = 0..3 from gx) of one image
// row, with the image's constants q (im_q row) and centre n: es (s channel, Y00 folded in) and
// ep (p channel over r, sqrt(3/4pi) folded in; beta_1m = ep * (dy, dz, dx)).  Points outside
// [u0, u0 + np) get zeros.  Branch-free, so the four points' exponentials and square roots
// overlap.  PF 2: one exponent for both channels.
template <class T, int PF>
__device__ __forceinline__ void sp_values4(const T (&q)[8], int nx0, int gx, int u0, int np, T hx, T ryz, T (&dx)[4],
                                           T (&es)[4], T (&ep)[4]) {
    const T dx0 = hx * ((T)(gx - nx0) - q[0]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        dx[u] = dx0 + (T)u * hx;
        const T rr = dx[u] * dx[u] + ryz;
        const T ir = (T(1.0e-14) > rr) ? T(0) : inv_sqrt<T>(rr);
        T s, p;
        if constexpr (PF == 2) {
            const T e = gauss<T>(rr, q[4]);
            s = q[3] * e;
            p = q[5] * e * ir;
        } else {
            s = q[3] * gauss<T>(rr, q[4]);
            p = q[5] * gauss<T>(rr, q[6]) * ir;
        }
        const bool on = (unsigned)(u - u0) < (unsigned)np;
        es[u] = on ? s : T(0);
        ep[u] = on ? p : T(0);
    }
}

template <class T>
__device__ __forceinline__ void load_q(const ProjLds<T>& L, int li, T (&q)[8]) {
    const Row<T>* src = reinterpret_cast<const Row<T>*>(L.im_q + 8 * li);
    constexpr int VN = 16 / sizeof(T);
#pragma unroll
    for (int k = 0; k < 8 / VN; ++k) {
        const Row<T> w = src[k];
#pragma unroll
        for (int u = 0; u < VN; ++u) q[k * VN + u] = w.v[u];
    }
}

// beta_j at point (ix, iy, iz) of column image li, for coefficient slots j < NJS (zero for
// slots of projectors the image does not hold).  Projector form PF: 0 general; 1 the image
// holds exactly an s channel in slot 0 and a p channel in slots 1..3; 2 as 1 with one Gaussian
// exponent for both channels (one exponential per point).
template <class T, int NJS, int PF>
__device__ __forceinline__ void beta_at(const ProjLds<T>& L, int li, T dx, T dy, T dz, T (&beta)[NJS]) {
    const T rr = dx * dx + (dy * dy + dz * dz);
    T ir = T(0);
    if (!(T(1.0e-14) > rr)) ir = inv_sqrt<T>(rr);
    const int* ii = L.im_i + 5 * li;
    if constexpr (PF != 0) {
        const int pb = ii[3];
        T es, ep;
        if constexpr (PF == 2) {
            const T e = gauss<T>(rr, L.pr_f[2 * pb + 1]);
            es = L.pr_f[2 * pb] * e;
            ep = L.pr_f[2 * pb + 2] * e * ir;
        } else {
            es = L.pr_f[2 * pb] * gauss<T>(rr, L.pr_f[2 * pb + 1]);
            ep = L.pr_f[2 * pb + 2] * gauss<T>(rr, L.pr_f[2 * pb + 3]) * ir;
        }
        beta[0] = es;
        beta[1] = ep * dy;
        beta[2] = ep * dz;
        beta[3] = ep * dx;
    } else {
#pragma unroll
        for (int j = 0; j < NJS; ++j) beta[j] = T(0);
        for (int p = ii[3]; p < ii[4]; ++p) {
            const int l = L.pr_i[2 * p], cr = L.pr_i[2 * p + 1];
            const T rad = L.pr_f[2 * p] * gauss<T>(rr, L.pr_f[2 * p + 1]);
            if (l == 0) {
#pragma unroll
                for (int j = 0; j < NJS; ++j)
                    if (j == cr) beta[j] = rad;
            } else {
                const T rp = rad * ir;
#pragma unroll
                for (int j = 0; j < NJS; ++j) {
                    if (j == cr) beta[j] = rp * dy;
                    if (j == cr + 1) beta[j] = rp * dz;
                    if (j == cr + 2) beta[j] = rp * dx;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------ projection ----
constexpr int kProjThreads = kRowsProjThreads;
constexpr int kProjPV = 8;      // staged psi vectors per thread per plane
constexpr int kItemsPT = 4;     // staged items per thread per plane (both kernels)
constexpr int kGroupsPT = 2;    // staged group run offsets per thread per plane (apply kernel)

// LDS offset rounded up to 16 bytes
__host__ __device__ constexpr size_t align16(size_t b) { return (b + 15) & ~size_t(15); }

template <class T, int NBP, int NJS, int PF>
__global__ __launch_bounds__(kProjThreads) void rows_project_kernel(RowsProjArgs a, const T* __restrict__ psi) {
    constexpr int VN = 16 / sizeof(T);
    extern __shared__ __align__(16) unsigned char smem[];
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int wg = xcd_order(blockIdx.x, gridDim.x);
    const int tile = wg % a.ntile;
    const int rest = wg / a.ntile;
    const int slab = rest % a.nslab;
    const int b0 = (rest / a.nslab) * NBP;
    const int y0 = tile * a.ty;
    const int rows = min(a.ty, ny - y0);
    const int z0 = slab * a.tz, z1 = min(z0 + a.tz, nz);
    const int tid = threadIdx.x;
    const int plane = nx * ny;

    // LDS: the tile plane of NBP bands, band fastest [row][x][band], reused for the reduction;
    // the plane's items; the slab's per-plane item offsets; the projector staging
    const size_t pl_elems = (size_t)a.ty * nx * NBP;
    const size_t red_elems = (size_t)kProjThreads * NJS * NBP;
    T* pl = reinterpret_cast<T*>(smem);
    size_t off = align16(sizeof(T) * (pl_elems > red_elems ? pl_elems : red_elems));
    uint32_t* itl = reinterpret_cast<uint32_t*>(smem + off);
    off += sizeof(uint32_t) * kProjThreads * kItemsPT;
    int* itoffl = reinterpret_cast<int*>(smem + off);
    off = align16(off + sizeof(int) * (a.tz + 1));
    const ProjLds<T> L = carve_proj<T, PF == 0>(smem + off, a.max_col, a.nproj);

    const int c0 = a.col_off[tile];
    const int ncol = a.col_off[tile + 1] - c0;
    stage_proj<T, PF == 0>(L, a.images, a.projs, a.nproj, a.col_img + c0, ncol, a.h, tid, kProjThreads);
    for (int i = tid; i <= z1 - z0; i += kProjThreads) itoffl[i] = a.it_off[(size_t)tile * nz + z0 + i];

    const uint32_t tm = a.thr_map[((size_t)tile * a.nslab + slab) * kProjThreads + tid];
    const bool act = tm != 0xffffffffu;
    const int li = (int)(tm & 1023u);
    const int rk = (int)((tm >> 10) & 255u);
    const int cnt = (int)((tm >> 18) & 255u) + 1;

    // staging assignment: vector v -> band v % NBP, then x-vector, then row; NBP divides the
    // thread count, so every vector of a thread belongs to band tid % NBP
    static_assert(kProjThreads % NBP == 0, "bands per workgroup must divide the thread count");
    const int nq = nx / VN;
    const int nvec = NBP * rows * nq;
    const int bs = tid % NBP;
    // lanes of a wavefront read different bands: plain global loads (a buffer resource must
    // be wavefront-uniform)
    const Row<T>* src = reinterpret_cast<const Row<T>*>(psi + (int64_t)min(b0 + bs, a.nbands - 1) * a.ld);
    int vo[kProjPV], lo[kProjPV];
#pragma unroll
    for (int k = 0; k < kProjPV; ++k) {
        const int v = tid + kProjThreads * k;
        const bool ok = v < nvec;
        const int rr = ok ? v / NBP : 0;
        const int xq = rr % nq, r = rr / nq;
        vo[k] = ok ? ((y0 + r) * nx + xq * VN) / VN : 0;   // in vectors
        lo[k] = ok ? ((r * nx + xq * VN) * NBP + bs) : -1;
    }
    const int* plb = a.pl_off + a.pl_base[tile];
    // next plane, held in registers until it is staged: psi vectors, items (dealt over the
    // workgroup), and the item range of this thread's image.  No global load is waited for
    // inside the compute phase, so these loads stay in flight across it.
    Row<T> pre[kProjPV];
    uint32_t ipre[kItemsPT];
    int nxt_s = 0, nxt_e = 0;
    // every load is issued by every lane (indices clamped into the arrays; unused values are
    // dropped), so the compiler waits for exactly the loads it needs
    const int lim = max(a.nitems - 1, 0);
    const int pli0 = act ? li : 0, pli1 = act ? li + 1 : 0;
    auto load = [&](int z) {
        const Row<T>* sp = src + (int64_t)z * (plane / VN);
#pragma unroll
        for (int k = 0; k < kProjPV; ++k) pre[k] = sp[vo[k]];
        const int ib = itoffl[z - z0];
#pragma unroll
        for (int k = 0; k < kItemsPT; ++k) ipre[k] = a.items[min(ib + tid + kProjThreads * k, lim)];
        const int* po = plb + (size_t)z * (ncol + 1);
        nxt_s = po[pli0];
        nxt_e = po[pli1];
    };

    T acc[NJS][NBP];
#pragma unroll
    for (int j = 0; j < NJS; ++j)
#pragma unroll
        for (int b = 0; b < NBP; ++b) acc[j][b] = T(0);

    const T hx = (T)a.h[0], hy = (T)a.h[1], hz = (T)a.h[2];
    __syncthreads();   // itoffl and the projector staging
    // s + p forms: this thread's image (centre and projector constants) in registers for the slab
    T iq[8];
    int in0 = 0, in1 = 0, in2 = 0;
    if constexpr (PF != 0) {
        load_q<T>(L, act ? li : 0, iq);
        in0 = L.im_n[4 * (act ? li : 0)];
        in1 = L.im_n[4 * (act ? li : 0) + 1];
        in2 = L.im_n[4 * (act ? li : 0) + 2];
    }
    load(z0);
    for (int z = z0; z < z1; ++z) {
        __syncthreads();   // the previous plane's readers are done
#pragma unroll
        for (int k = 0; k < kProjPV; ++k)
            if (lo[k] >= 0) {
#pragma unroll
                for (int i = 0; i < VN; ++i) pl[lo[k] + i * NBP] = pre[k].v[i];
            }
        {
            const int n = itoffl[z - z0 + 1] - itoffl[z - z0];
#pragma unroll
            for (int k = 0; k < kItemsPT; ++k) {
                const int i = tid + kProjThreads * k;
                if (i < n) itl[i] = ipre[k];
            }
        }
        const int cs = nxt_s, ce = nxt_e;
        if (z + 1 < z1) load(z + 1);
        __syncthreads();
        if (PF != 0 && act) {
            const T dz = hz * ((T)(z - in2) - iq[2]);
            const T dzz = dz * dz;
            for (int i = cs + rk; i < ce; i += cnt) {
                const uint32_t it = itl[i];
                const int r = (int)(it & 31u), xs = (int)((it >> 5) & 1023u), np = (int)((it >> 15) & 3u) + 1;
                const T dy = hy * ((T)(y0 + r - in1) - iq[1]);
                const T* prow = pl + (size_t)(r * nx + xs) * NBP;
                T dx[4], es[4], ep[4];
                sp_values4<T, PF>(iq, in0, xs, 0, np, hx, dy * dy + dzz, dx, es, ep);
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    const T beta[4] = {es[p], ep[p] * dy, ep[p] * dz, ep[p] * dx[p]};
                    // points past the item read its first point (in the row) with zero weights
                    const Row<T>* pv = reinterpret_cast<const Row<T>*>(prow + (p < np ? p : 0) * NBP);
#pragma unroll
                    for (int q = 0; q < NBP / VN; ++q) {
                        const Row<T> w = pv[q];
#pragma unroll
                        for (int u = 0; u < VN; ++u)
#pragma unroll
                            for (int j = 0; j < 4; ++j) acc[j][q * VN + u] += beta[j] * w.v[u];
                    }
                }
            }
        } else if (act) {
            const int* ii = L.im_i + 5 * li;
            const T* ff = L.im_f + 3 * li;
            const T dz = hz * ((T)(z - ii[2]) - ff[2]);
            for (int i = cs + rk; i < ce; i += cnt) {
                const uint32_t it = itl[i];
                const int r = (int)(it & 31u), xs = (int)((it >> 5) & 1023u), np = (int)((it >> 15) & 3u) + 1;
                const T dy = hy * ((T)(y0 + r - ii[1]) - ff[1]);
                const T* prow = pl + (size_t)(r * nx + xs) * NBP;
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    if (p < np) {
                        const T dx = hx * ((T)(xs + p - ii[0]) - ff[0]);
                        T beta[NJS];
                        beta_at<T, NJS, PF>(L, li, dx, dy, dz, beta);
                        const Row<T>* pv = reinterpret_cast<const Row<T>*>(prow + p * NBP);
#pragma unroll
                        for (int q = 0; q < NBP / VN; ++q) {
                            const Row<T> w = pv[q];
#pragma unroll
                            for (int u = 0; u < VN; ++u)
#pragma unroll
                                for (int j = 0; j < NJS; ++j) acc[j][q * VN + u] += beta[j] * w.v[u];
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    T* red = pl;
#pragma unroll
    for (int j = 0; j < NJS; ++j)
#pragma unroll
        for (int b = 0; b < NBP; ++b) red[(size_t)tid * NJS * NBP + j * NBP + b] = acc[j][b];
    __syncthreads();
    if (act) {
        const int tb = tid - rk;
        const int slot = a.thr_slot[((size_t)tile * a.nslab + slab) * kProjThreads + tid];
        for (int o = rk; o < NJS * NBP; o += cnt) {
            double sum = 0.0;
            for (int u = 0; u < cnt; ++u) sum += (double)red[(size_t)(tb + u) * NJS * NBP + o];
            const int j = o / NBP, b = o - j * NBP;
            if (b0 + b < a.nbands) a.part[((size_t)slot * NJS + j) * a.part_ld + b0 + b] = sum;
        }
    }
}

// coef[b * ncoef + off(a) + j] = dV * sum of the atom's partials
__global__ void rows_coef_reduce_kernel(const double* __restrict__ part, int part_ld, int njs,
                                        const int* __restrict__ atom_slot_off, const int* __restrict__ atom_slot,
                                        const int* __restrict__ atom_coef_off, int natoms, int nbands, double dv,
                                        double* __restrict__ coef, int ncoef) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)natoms * njs * nbands) return;
    const int b = (int)(idx % nbands);
    const int64_t rest = idx / nbands;
    const int j = (int)(rest % njs);
    const int at = (int)(rest / njs);
    const int cb = atom_coef_off[at];
    if (j >= atom_coef_off[at + 1] - cb) return;
    double sum = 0.0;
    for (int s = atom_slot_off[at]; s < atom_slot_off[at + 1]; ++s)
        sum += part[((size_t)atom_slot[s] * njs + j) * part_ld + b];
    coef[(int64_t)b * ncoef + cb + j] = dv * sum;
}

// -------------------------------------------------------------- stencil + V + projector ----
constexpr int kRowsMaxThreads = 1024;
// buffer offset of lanes with nothing to load or store: beyond every band's num_records, so
// the hardware returns zero for the load and drops the store
constexpr int kOOB = 0x7ffffff0;

template <class T, int NBW, int NJS, int PF, bool NL>
__global__ __launch_bounds__(NBW >= 4 ? 1024 : 512) void rows_hpsi_kernel(RowsHpsiArgs a, Coef8<T> sc,
                                                                          const T* __restrict__ V,
                                                                          const T* __restrict__ psi,
                                                                          T* __restrict__ out) {
    constexpr int VN = 16 / sizeof(T);
    constexpr int S = sizeof(T);
    extern __shared__ __align__(16) unsigned char smem[];
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int W = nx + 8;
    const int ty = a.ty;
    const int lband = (ty + 8) * W;
    const int nthr = blockDim.x;
    // LDS layout (rows_hpsi_lds on the host): planes, per-item projector values of one chunk of
    // nthr items, E*c, the plane's items, group run offsets (two planes), the slab's per-plane
    // item offsets, projector staging
    const int ngo = ty * a.ng4;                                // groups of a tile plane
    T* planes = reinterpret_cast<T*>(smem);                    // [2][NBW][ty + 8][W]
    T* stg = planes + (size_t)2 * NBW * lband;                 // [nthr][NBW][4]
    T* ec = stg + (size_t)nthr * NBW * 4;                      // [max_col][NBW][NJS]
    size_t off = sizeof(T) * ((size_t)2 * NBW * lband + (size_t)nthr * NBW * 4 + (size_t)a.max_col * NBW * NJS);
    uint32_t* itl = reinterpret_cast<uint32_t*>(smem + off);   // [nthr * kItemsPT]
    off += sizeof(uint32_t) * (size_t)nthr * kItemsPT;
    int* gol = reinterpret_cast<int*>(smem + off);             // [2][ngo + 1]
    off += sizeof(int) * (size_t)2 * (ngo + 1);
    int* itoffl = reinterpret_cast<int*>(smem + off);          // [tz + 1]
    off = align16(off + sizeof(int) * (a.tz + 1));
    const ProjLds<T> L = carve_proj<T, PF == 0>(smem + off, a.max_col, a.nproj);

    const int wg = xcd_order(blockIdx.x, gridDim.x);
    const int tile = wg % a.ntile;
    const int rest = wg / a.ntile;
    const int slab = rest % a.nslab;
    const int b0 = (rest / a.nslab) * NBW;
    const int y0 = tile * ty;
    const int rows = min(ty, ny - y0);
    const int z0 = slab * a.tz, z1 = min(z0 + a.tz, nz);
    const int tid = threadIdx.x;
    // tpb is a multiple of 64: every lane of a wavefront works on the same band
    const int bl = __builtin_amdgcn_readfirstlane(tid / a.tpb);
    const int lt = tid - bl * a.tpb;
    const int nq = nx / VN;
    const int r = lt / nq, xq = lt - r * nq;
    const int x = xq * VN;
    const bool own = bl < NBW && r < rows;
    const bool bok = (b0 + bl) < a.nbands;
    const int plane = nx * ny;
    const uint32_t band_bytes = (uint32_t)((int64_t)plane * nz * S);
    const int band = min(b0 + min(bl, NBW - 1), a.nbands - 1);
    const rsrc_t rin = rsrc(psi + (int64_t)band * a.ld_in, band_bytes);
    const rsrc_t rout = rsrc(out + (int64_t)band * a.ld_out, band_bytes);
    // second operand of the point term: V (H psi), the previous output (accumulating
    // Laplacian), or nothing (a resource of zero records reads zeros)
    const rsrc_t rsec = a.add_v ? rsrc(V, band_bytes) : (a.accumulate ? rout : rsrc(psi, 0));

    // own vector: plane offset (kOOB for lanes without a point), LDS slot and the wrapped copy
    // that pads the x-halo
    const int vo_pt = ((y0 + r) * nx + x) * S;
    const int vo = own ? vo_pt : kOOB;
    const int vst = (own && bok) ? vo_pt : kOOB;
    const int lc = (r + 4) * W + 4 + x;
    const int lw = (x < 4) ? lc + nx : ((x >= nx - 4) ? lc - nx : -1);
    // halo vector of this thread: 8 rows x nq vectors per band (4 above the tile, 4 below)
    const bool hown = bl < NBW && lt < 8 * nq;
    int ho = kOOB, hl = 0, hw = -1;
    if (hown) {
        const int hr = lt / nq, hx = (lt - hr * nq) * VN;
        int gy = (hr < 4) ? y0 - 4 + hr : y0 + rows + (hr - 4);
        gy = ((gy % ny) + ny) % ny;
        const int lrow = (hr < 4) ? hr : rows + hr;
        ho = (gy * nx + hx) * S;
        hl = lrow * W + 4 + hx;
        hw = (hx < 4) ? hl + nx : ((hx >= nx - 4) ? hl - nx : -1);
    }
    auto zoff = [&](int z) {
        int zz = z % nz;
        zz = zz < 0 ? zz + nz : zz;
        return zz * plane * S;
    };

    // projector staging: image data of the tile column and E*c of its atoms for the bands, the
    // slab's per-plane item offsets
    if constexpr (NL) {
        const int c0 = a.col_off[tile];
        const int ncol = a.col_off[tile + 1] - c0;
        stage_proj<T, PF == 0>(L, a.images, a.projs, a.nproj, a.col_img + c0, ncol, a.h, tid, nthr);
        for (int i = tid; i < ncol * NBW * NJS; i += nthr) {
            const int li = i / (NBW * NJS), b = (i / NJS) % NBW, j = i % NJS;
            const int at = a.images[a.col_img[c0 + li]].atom;
            const int cb = a.atom_coef_off[at];
            T v = T(0);
            if (j < a.atom_coef_off[at + 1] - cb && b0 + b < a.nbands)
                v = (T)(a.coef_energy[cb + j] * a.coef[(int64_t)(b0 + b) * a.ncoef + cb + j]);
            ec[i] = v;
        }
        for (int i = tid; i <= z1 - z0; i += nthr) itoffl[i] = a.it_off[(size_t)tile * nz + z0 + i];
        __syncthreads();
    }

    Row<T> q[9];
#pragma unroll
    for (int d = 0; d < 8; ++d) q[d] = vload<T>(rin, vo, zoff(z0 - 4 + d));
    // the eight warm-up planes complete here, once: the loop's waits then cover only the
    // loads issued inside it
    __builtin_amdgcn_s_waitcnt(0);
    // next plane, held in registers until it is staged: psi (4 planes ahead in z), V or the
    // accumulated output, the y-halo, and the plane's projector items dealt over the
    // workgroup.  Every lane issues every load (out-of-range offsets or clamped indices where
    // it has nothing to load), so the compiler can wait for exactly the loads it needs, and
    // the projector phase reads LDS only: these loads stay in flight across the plane's work.
    Row<T> qn{}, hn{}, vn{};
    uint32_t ipre[kItemsPT];
    int gpre[kGroupsPT];
    // items and group offsets through buffer resources: a load past the end of the item array
    // returns zero (such slots are never read), and each offset is one add to a per-thread
    // constant
    const rsrc_t ritems = rsrc(a.items, (uint32_t)((int64_t)a.nitems * 4));
    const rsrc_t rgrp = rsrc(a.grp_off, (uint32_t)((int64_t)a.ntile * nz * (ngo + 1) * 4));
    auto prefetch = [&](int z) {
        qn = vload<T>(rin, vo, zoff(z + 4));
        vn = vload<T>(rsec, vo, zoff(z));
        hn = vload<T>(rin, ho, zoff(z));
        if constexpr (NL) {
            const int ib4 = __builtin_amdgcn_readfirstlane(itoffl[z - z0]) * 4;
#pragma unroll
            for (int k = 0; k < kItemsPT; ++k)
                ipre[k] = __builtin_amdgcn_raw_buffer_load_b32(ritems, ib4 + (tid + nthr * k) * 4, 0, 0);
            const int gb4 = (tile * nz + z) * (ngo + 1) * 4;
#pragma unroll
            for (int k = 0; k < kGroupsPT; ++k)
                gpre[k] = (int)__builtin_amdgcn_raw_buffer_load_b32(rgrp, gb4 + min(tid + nthr * k, ngo) * 4, 0, 0);
        }
    };
    prefetch(z0);
    const T hx = (T)a.h[0], hy = (T)a.h[1], hz = (T)a.h[2];

    // projector values of item i of the plane at z into per-item slot `slot`: the item (points
    // of one image in one 4-point group) is evaluated once for the workgroup's bands; no two
    // threads write one slot (no atomics)
    auto eval_item = [&](int i, int slot, int z) {
        const uint32_t it = itl[i];
        const int ir = (int)(it & 31u), xs = (int)((it >> 5) & 1023u);
        const int np = (int)((it >> 15) & 3u) + 1, li = (int)(it >> 17);
        T val[NBW][4];
        if constexpr (PF != 0) {
            // the item's image constants and the E*c of its atom: independent LDS reads
            const int4 cn = *reinterpret_cast<const int4*>(L.im_n + 4 * li);
            T cq[8];
            load_q<T>(L, li, cq);
            T e[NBW][4];
#pragma unroll
            for (int b = 0; b < NBW; ++b) {
                const Row<T>* ep = reinterpret_cast<const Row<T>*>(ec + ((size_t)li * NBW + b) * NJS);
#pragma unroll
                for (int q4 = 0; q4 < 4 / VN; ++q4) {
                    const Row<T> w = ep[q4];
#pragma unroll
                    for (int u = 0; u < VN; ++u) e[b][q4 * VN + u] = w.v[u];
                }
            }
            const T dy = hy * ((T)(y0 + ir - cn.y) - cq[1]);
            const T dz = hz * ((T)(z - cn.z) - cq[2]);
            T dx[4], es[4], ep[4];
            sp_values4<T, PF>(cq, cn.x, xs & ~3, xs & 3, np, hx, dy * dy + dz * dz, dx, es, ep);
#pragma unroll
            for (int b = 0; b < NBW; ++b) {
                const T ayz = e[b][1] * dy + e[b][2] * dz;
#pragma unroll
                for (int u = 0; u < 4; ++u) val[b][u] = e[b][0] * es[u] + ep[u] * (ayz + e[b][3] * dx[u]);
            }
        } else {
            const int* ii = L.im_i + 5 * li;
            const T* ff = L.im_f + 3 * li;
            const T dy = hy * ((T)(y0 + ir - ii[1]) - ff[1]);
            const T dz = hz * ((T)(z - ii[2]) - ff[2]);
            T e[NBW][NJS];
#pragma unroll
            for (int b = 0; b < NBW; ++b) {
                const Row<T>* ep = reinterpret_cast<const Row<T>*>(ec + ((size_t)li * NBW + b) * NJS);
#pragma unroll
                for (int q4 = 0; q4 < NJS / VN; ++q4) {
                    const Row<T> w = ep[q4];
#pragma unroll
                    for (int u = 0; u < VN; ++u) e[b][q4 * VN + u] = w.v[u];
                }
            }
            const int gx = xs & ~3, u0 = xs & 3;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
#pragma unroll
                for (int b = 0; b < NBW; ++b) val[b][u] = T(0);
                if (u >= u0 && u < u0 + np) {
                    const T dx = hx * ((T)(gx + u - ii[0]) - ff[0]);
                    T beta[NJS];
                    beta_at<T, NJS, PF>(L, li, dx, dy, dz, beta);
#pragma unroll
                    for (int b = 0; b < NBW; ++b) {
                        T v = T(0);
#pragma unroll
                        for (int j = 0; j < NJS; ++j) v += e[b][j] * beta[j];
                        val[b][u] = v;
                    }
                }
            }
        }
#pragma unroll
        for (int b = 0; b < NBW; ++b) {
            Row<T>* dst = reinterpret_cast<Row<T>*>(stg + ((size_t)slot * NBW + b) * 4);
#pragma unroll
            for (int h = 0; h < 4 / VN; ++h) {
                Row<T> w;
#pragma unroll
                for (int u = 0; u < VN; ++u) w.v[u] = val[b][h * VN + u];
                dst[h] = w;
            }
        }
    };

    // a plane's result is stored at the top of the next iteration, after that iteration's
    // loads were issued one iteration earlier: every iteration starts with the same memory
    // operations outstanding (the prefetch, then one store), so waiting for the prefetch leaves
    // the store in flight.  The first iteration's store is a dropped out-of-range one.
    Row<T> res{};
    int st_off = kOOB, st_z = 0;

    // q[k] is the plane z - 4 + k of the z-queue
    for (int z = z0; z < z1; ++z) {
        const int buf = (z - z0) & 1;
        T* Lb = planes + ((size_t)buf * NBW + (bl < NBW ? bl : 0)) * lband;
        int* golb = gol + (size_t)buf * (ngo + 1);
        vstore<T>(rout, res, st_off, st_z);
        q[8] = qn;
        const Row<T> vcur = vn;
        if (own) {
            *reinterpret_cast<Row<T>*>(Lb + lc) = q[4];
            if (lw >= 0) *reinterpret_cast<Row<T>*>(Lb + lw) = q[4];
        }
        if (hown) {
            *reinterpret_cast<Row<T>*>(Lb + hl) = hn;
            if (hw >= 0) *reinterpret_cast<Row<T>*>(Lb + hw) = hn;
        }
        int nit = 0;
        if constexpr (NL) {
            // this plane's items (their readers of the previous plane passed that plane's last
            // chunk barrier) and group runs (a buffer per plane parity: the previous plane's
            // readers may still be gathering)
            nit = itoffl[z - z0 + 1] - itoffl[z - z0];
            // all nthr * kItemsPT slots (those past the plane's items are never read)
#pragma unroll
            for (int k = 0; k < kItemsPT; ++k) itl[tid + nthr * k] = ipre[k];
#pragma unroll
            for (int k = 0; k < kGroupsPT; ++k) {
                const int i = tid + nthr * k;
                if (nthr * k <= ngo && i <= ngo) golb[i] = gpre[k];
            }
        }
        if (z + 1 < z1) prefetch(z + 1);
        __syncthreads();

        if (own) {
            // V psi and the stencil: centre and z from the register queue, x and y from LDS,
            // accumulated term by term so few values are live at once
            const T* Lc = Lb + (r + 4) * W + x;   // Lc[4 + i] is point x + i
            T d[VN];
#pragma unroll
            for (int i = 0; i < VN; ++i) {
                d[i] = sc.c0 * q[4].v[i];
                d[i] += sc.cz[3] * (q[8].v[i] + q[0].v[i]) + sc.cz[2] * (q[7].v[i] + q[1].v[i]) +
                        sc.cz[1] * (q[6].v[i] + q[2].v[i]) + sc.cz[0] * (q[5].v[i] + q[3].v[i]);
            }
            {
                T xsv[VN + 8];
#pragma unroll
                for (int k = 0; k < (VN + 8) / VN; ++k) {
                    const Row<T> w = *reinterpret_cast<const Row<T>*>(Lc + k * VN);
#pragma unroll
                    for (int u = 0; u < VN; ++u) xsv[k * VN + u] = w.v[u];
                }
#pragma unroll
                for (int i = 0; i < VN; ++i)
                    d[i] += sc.cx[3] * (xsv[i + 8] + xsv[i]) + sc.cx[2] * (xsv[i + 7] + xsv[i + 1]) +
                            sc.cx[1] * (xsv[i + 6] + xsv[i + 2]) + sc.cx[0] * (xsv[i + 5] + xsv[i + 3]);
            }
#pragma unroll
            for (int m = 3; m >= 0; --m) {
                const Row<T> ym = *reinterpret_cast<const Row<T>*>(Lc + 4 - (m + 1) * W);
                const Row<T> yp = *reinterpret_cast<const Row<T>*>(Lc + 4 + (m + 1) * W);
#pragma unroll
                for (int i = 0; i < VN; ++i) d[i] += sc.cy[m] * (yp.v[i] + ym.v[i]);
            }
#pragma unroll
            for (int i = 0; i < VN; ++i) res.v[i] = (a.add_v ? vcur.v[i] * q[4].v[i] : vcur.v[i]) + d[i];
        }

        if constexpr (NL) {
            // projector term in chunks of nthr items (one item per thread per chunk): values
            // into the chunk's slots, then each point adds the values of the items of its
            // 4-point group's run that lie in the chunk
            int s0 = 0, s1 = 0;
            if (own) {
                const int gi = r * a.ng4 + (x >> 2);
                s0 = golb[gi];
                s1 = golb[gi + 1];
            }
            const int hsel = (VN == 4) ? 0 : ((x >> 1) & 1);   // FP64: which half of the group
            for (int c0 = 0; c0 < nit; c0 += nthr) {
                if (c0 + tid < nit) eval_item(c0 + tid, tid, z);
                __syncthreads();
                if (own) {
                    const int e1 = min(s1, c0 + nthr);
                    for (int i = max(s0, c0); i < e1; ++i) {
                        const Row<T> w = reinterpret_cast<const Row<T>*>(stg + ((size_t)(i - c0) * NBW + bl) * 4)[hsel];
#pragma unroll
                        for (int u = 0; u < VN; ++u) res.v[u] += w.v[u];
                    }
                }
                if (c0 + nthr < nit) __syncthreads();   // the next chunk reuses the slots
            }
        }
        // every lane stores (next iteration): lanes without a point, or past the last band, use kOOB
        st_off = vst;
        st_z = zoff(z);
#pragma unroll
        for (int d = 0; d < 8; ++d) q[d] = q[d + 1];
    }
    vstore<T>(rout, res, st_off, st_z);
}

// ------------------------------------------------------------------------ host side ----
namespace {

template <class T>
Coef8<T> coef8(const qg_grid& g, double coef) {
    const double w[4] = {8064.0, -1008.0, 128.0, -9.0};
    Coef8<T> s;
    s.c0 = (T)(-14350.0 / 5040.0 * coef *
               (1.0 / (g.h[0] * g.h[0]) + 1.0 / (g.h[1] * g.h[1]) + 1.0 / (g.h[2] * g.h[2])));
    for (int m = 0; m < 4; ++m) {
        s.cx[m] = (T)(w[m] * coef / (5040.0 * g.h[0] * g.h[0]));
        s.cy[m] = (T)(w[m] * coef / (5040.0 * g.h[1] * g.h[1]));
        s.cz[m] = (T)(w[m] * coef / (5040.0 * g.h[2] * g.h[2]));
    }
    return s;
}

int round64i(int v) { return (v + 63) / 64 * 64; }

}  // namespace

// Tile height of the apply kernel for a row of nq vectors: whole wavefronts per band with as
// few idle lanes as possible (rows of 8-16 preferred), at least 8 rows' worth of threads for
// the y-halo, and at most kRowsBandThreads threads per band (the kernel's launch bound with one
// band per workgroup).  0 when even 8 rows exceed that.
constexpr int kRowsBandThreads = 512;
static int rows_pick_ty(int nq, int ny, int want) {
    if (round64i(8 * nq) > kRowsBandThreads) return 0;
    if (want > 0) {
        int ty = std::max(1, std::min(ny, std::min(32, want)));
        while (ty > 8 && round64i(ty * nq) > kRowsBandThreads) --ty;
        return ty;
    }
    // 8 rows (more workgroups, less LDS: measured best at 80^3 and 120^3) unless that leaves
    // more than a quarter of the lanes idle
    if (8 <= ny && 4 * 8 * nq >= 3 * round64i(8 * nq)) return 8;
    int best = 8, bests = -1;
    for (int ty : {8, 16, 12, 10, 6, 20, 24, 4, 32}) {
        if (ty > ny) continue;
        const int t = round64i(std::max(ty, 8) * nq);
        if (t > kRowsBandThreads) continue;
        const int score = (int)(1000.0 * ty * nq / t) + (ty >= 8 ? 5 : 0);
        if (score > bests) { bests = score; best = ty; }
    }
    return std::min(best, ny);
}

// Tile height of the projection kernel: the largest (at most 32 rows, at most ny) whose plane
// of nbp bands fits the staging registers (kProjPV vectors per thread); 0 when one row does not.
static int rows_pick_typ(int nq, int ny, int nbp) {
    const int cap = kProjThreads * kProjPV / std::max(1, nbp * nq);
    return std::max(0, std::min(std::min(32, ny), cap));
}

// Whether the row kernels serve this grid and projector set: x-rows of whole 16-byte vectors,
// Gaussian projectors with l <= 1 and at most 8 coefficient slots per atom.
static bool rows_eligible(const qg_grid& g, int dtype, const qg_nonlocal* nl) {
    if (env_int("QG_ROWS", 1) == 0) return false;
    const int VN = dtype == QG_F32 ? 4 : 2;
    if (g.n[0] % VN != 0 || g.n[0] < 8 || g.n[0] > 1024) return false;
    if (g.ng() * 8 > (int64_t)kOOB - 64) return false;   // band bytes below the out-of-range offset
    if (rows_pick_ty(g.n[0] / VN, g.n[1], 0) == 0) return false;
    if (nl) {
        if (nl->geo.max_coef_per_atom > 8) return false;
        for (const DevProj& p : nl->geo.projs)
            if (p.kind != QG_RADIAL_GAUSSIAN || p.l > 1) return false;
    }
    return true;
}

static int rows_nthr(const qg_grid& g, int dtype, int ty, int nbw) {
    const int VN = dtype == QG_F32 ? 4 : 2;
    return nbw * round64i(std::max(ty, 8) * (g.n[0] / VN));
}

struct RowsShape {
    int ty, tz;     // apply kernel: tile rows, slab planes
    int typ, tzp;   // projection kernel: tile rows, slab planes
    int nbw, nbp;   // bands per workgroup of the apply and projection kernels
};

// Launch shape for nbands bands (0: only the geometry-defining fields, ty and tzp, which do not
// depend on the band count).  QG_ROWS_* environment variables override (measurement switches).
static RowsShape rows_shape(const qg_grid& g, int dtype, int nbands, const qg_hpsi_plan* plan = nullptr) {
    const int VN = dtype == QG_F32 ? 4 : 2;
    const int nq = g.n[0] / VN;
    const qg_hpsi_plan pl = plan ? *plan : qg_hpsi_plan{};
    RowsShape sh{};
    sh.ty = rows_pick_ty(nq, g.n[1], env_int("QG_ROWS_TY", pl.rows_ty));
    const int ntile = std::max(1, (g.n[1] + std::max(1, sh.ty) - 1) / std::max(1, sh.ty));
    // about 4096 workgroups per launch: 2 bands per workgroup (FP32) when the tiles and bands
    // give that many, else 1, and z-slabs (at least 8 planes each) for the rest
    constexpr int kWantWg = 4096;
    int nbw = 1;
    if (dtype == QG_F32 && (int64_t)ntile * ((nbands + 1) / 2) >= kWantWg / 2) nbw = 2;
    nbw = env_int("QG_ROWS_NBW", pl.rows_bands > 0 ? pl.rows_bands : nbw);
    if (nbw != 1 && nbw != 2 && nbw != 4) nbw = 1;
    while (nbw > 1 && nbw * round64i(std::max(sh.ty, 8) * nq) > (nbw >= 4 ? 1024 : 512)) nbw /= 2;
    sh.nbw = nbw;
    const int64_t wg = (int64_t)ntile * std::max(1, (nbands + nbw - 1) / nbw);
    const int nslab = (int)std::max<int64_t>(1, std::min<int64_t>((kWantWg + wg - 1) / wg, std::max(1, g.n[2] / 8)));
    sh.tz = env_int("QG_ROWS_TZ", pl.rows_tz > 0 ? pl.rows_tz : (g.n[2] + nslab - 1) / nslab);
    sh.tz = std::max(1, std::min(g.n[2], sh.tz));
    sh.tzp = std::max(1, std::min(g.n[2], env_int("QG_ROWS_TZP", pl.rows_tzp > 0 ? pl.rows_tzp : 8)));
    // FP64 projection with 4 bands per workgroup: its reduction buffer (256 threads x 4 slots
    // x bands) stays at 32 KB
    sh.nbp = env_int("QG_ROWS_NBP", pl.rows_nbp > 0 ? pl.rows_nbp : (dtype == QG_F32 ? 8 : 4)) == 4 ? 4 : 8;
    sh.typ = rows_pick_typ(nq, g.n[1], sh.nbp);
    return sh;
}

// The row geometry of nl for tile rows ty, projection slabs of tzp planes and item order
// (by 4-point group for the apply kernel, by image for the projection), built (and uploaded)
// on first use; with build == false only a cached one is returned (nullptr if none).
static RowsGeo* rows_geometry(qg_nonlocal* nl, int ty, int tzp, bool by_group, bool build = true) {
    std::lock_guard<std::mutex> lk(nl->rows_mu);
    for (auto& c : nl->rows)
        if (c && c->ty == ty && c->tzp == tzp && c->by_group == by_group) return c.get();
    if (!build) return nullptr;
    auto c = std::make_unique<RowsGeo>();
    if (!build_rows_geometry(nl->gcopy, nl->geo, ty, tzp, by_group, *c)) {
        c->ok = false;
    } else {
        c->ok = true;
        auto up = [](auto& buf, auto& vec) {
            if (vec.empty()) vec.push_back({});
            buf.upload(vec.data(), vec.size());
        };
        up(c->d_col_off, c->col_off);
        up(c->d_col_img, c->col_img);
        up(c->d_it_off, c->it_off);
        up(c->d_items, c->items);
        up(c->d_pl_base, c->pl_base);
        up(c->d_pl_off, c->pl_off);
        up(c->d_thr_map, c->thr_map);
        up(c->d_thr_slot, c->thr_slot);
        up(c->d_atom_slot_off, c->atom_slot_off);
        up(c->d_atom_slot, c->atom_slot);
        up(c->d_grp_off, c->grp_off);
        std::vector<uint32_t>().swap(c->items);
        std::vector<int>().swap(c->pl_off);
        std::vector<uint32_t>().swap(c->thr_map);
        std::vector<int>().swap(c->thr_slot);
        std::vector<int>().swap(c->grp_off);
    }
    nl->rows.push_back(std::move(c));
    return nl->rows.back().get();
}

// Most LDS one workgroup may allocate on the current device (160 KB on MI355X), looked up once
// per device.
static size_t lds_limit() {
    static std::mutex mu;
    static size_t cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
    std::lock_guard<std::mutex> lk(mu);
    if (!cached[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 65536;
        cached[dev] = (size_t)v;
    }
    return cached[dev];
}

// LDS bytes of a projection workgroup (layout of rows_project_kernel)
template <class T>
static size_t rows_proj_lds(const RowsGeo& rg, const qg_grid& g, int nbp, int nproj) {
    const size_t pl = (size_t)rg.ty * g.n[0] * nbp, red = (size_t)kProjThreads * rg.njs * nbp;
    size_t off = align16(sizeof(T) * std::max(pl, red));
    off += sizeof(uint32_t) * kProjThreads * kItemsPT;
    off = align16(off + sizeof(int) * (rg.tzp + 1));
    return off + proj_lds_bytes<T>(std::max(1, rg.max_col), nproj, rg.pform == 0);
}

// LDS bytes of an apply workgroup of nthr threads (layout of rows_hpsi_kernel): per-item values
// of one chunk of nthr items, ng4 groups per row, projector form pform
template <class T>
static size_t rows_hpsi_lds(const qg_grid& g, int ty, int tz, int nbw, int njs, int max_col, int nproj, int nthr,
                            int ng4, int pform) {
    const size_t W = g.n[0] + 8;
    size_t off = sizeof(T) * (2 * (size_t)nbw * (ty + 8) * W + (size_t)nthr * nbw * 4 + (size_t)max_col * nbw * njs);
    off += sizeof(uint32_t) * (size_t)nthr * kItemsPT;
    off += sizeof(int) * 2 * (size_t)(ty * ng4 + 1);
    off = align16(off + sizeof(int) * (tz + 1));
    return off + proj_lds_bytes<T>(max_col, nproj, pform == 0);
}

// Scratch of one application: projector partials and coefficients (grown, never shrunk).
template <class T, int NJS, int PF>
static void rows_launch_project(qg_nonlocal* nl, const RowsGeo& rg, int nbp, int nbands, const T* psi, int64_t ld,
                                double* part, double* coef, hipStream_t s) {
    const qg_grid& g = nl->gcopy;
    RowsProjArgs a{};
    a.nx = g.n[0]; a.ny = g.n[1]; a.nz = g.n[2];
    a.ty = rg.ty; a.tz = rg.tzp; a.ntile = rg.ntile; a.nslab = rg.nslabp; a.nbands = nbands;
    a.ld = ld;
    a.col_off = rg.d_col_off.p; a.col_img = rg.d_col_img.p;
    a.it_off = rg.d_it_off.p; a.items = rg.d_items.p;
    a.nitems = rg.it_off.empty() ? 0 : rg.it_off.back();
    a.pl_base = rg.d_pl_base.p; a.pl_off = rg.d_pl_off.p;
    a.thr_map = rg.d_thr_map.p; a.thr_slot = rg.d_thr_slot.p;
    a.images = nl->d_images.p; a.projs = nl->d_projs.p; a.nproj = (int)nl->geo.projs.size();
    a.part = part;
    a.part_ld = nbands;
    a.max_col = std::max(1, rg.max_col);
    for (int d = 0; d < 3; ++d) a.h[d] = g.h[d];
    const int VN = 16 / sizeof(T);
    QG_REQUIRE((int64_t)nbp * rg.ty * (g.n[0] / VN) <= (int64_t)kProjThreads * kProjPV,
               "projection tile does not fit the staging registers");
    QG_REQUIRE(rg.max_items <= kProjThreads * kItemsPT, "a tile plane holds more projector items than staged");
    const size_t lds = rows_proj_lds<T>(rg, g, nbp, a.nproj);
    QG_REQUIRE(lds <= lds_limit(), "projection tile does not fit in LDS");
    const int64_t nwg = (int64_t)rg.ntile * rg.nslabp * ((nbands + nbp - 1) / nbp);
    QG_REQUIRE(nwg < (1LL << 31), "too many workgroups");
    if (nbp == 8)
        hipLaunchKernelGGL((rows_project_kernel<T, 8, NJS, PF>), dim3((unsigned)nwg), dim3(kProjThreads), lds, s, a, psi);
    else
        hipLaunchKernelGGL((rows_project_kernel<T, 4, NJS, PF>), dim3((unsigned)nwg), dim3(kProjThreads), lds, s, a, psi);
    QG_HIP(hipGetLastError());
    const int64_t nred = (int64_t)nl->geo.natoms * NJS * nbands;
    if (nred > 0) {
        hipLaunchKernelGGL(rows_coef_reduce_kernel, dim3((unsigned)((nred + 255) / 256)), dim3(256), 0, s, part, nbands,
                           NJS, rg.d_atom_slot_off.p, rg.d_atom_slot.p, nl->d_atom_coef_off.p, nl->geo.natoms, nbands,
                           g.dvol(), coef, nl->geo.ncoef);
        QG_HIP(hipGetLastError());
    }
}

template <class T, int NBW, int NJS, int PF>
static void rows_launch_hpsi_t(const qg_grid& g, const RowsShape& sh, const RowsGeo* rg, qg_nonlocal* nl, int nbands,
                               const T* psi, int64_t ld_in, T* out, int64_t ld_out, double lap_coef, int add_v,
                               int accumulate, const T* V, const double* coef, hipStream_t s) {
    constexpr int VN = 16 / sizeof(T);
    const int nq = g.n[0] / VN;
    RowsHpsiArgs a{};
    a.nx = g.n[0]; a.ny = g.n[1]; a.nz = g.n[2];
    a.ty = sh.ty; a.tz = sh.tz;
    a.ntile = (g.n[1] + sh.ty - 1) / sh.ty;
    a.nslab = (g.n[2] + sh.tz - 1) / sh.tz;
    a.nbands = nbands;
    a.tpb = round64i(std::max(sh.ty, 8) * nq);
    a.ld_in = ld_in; a.ld_out = ld_out;
    a.add_v = add_v; a.accumulate = accumulate;
    a.has_nl = rg != nullptr;
    a.njs = NJS;
    a.max_col = 1;
    if (rg) {
        a.col_off = rg->d_col_off.p; a.col_img = rg->d_col_img.p;
        a.it_off = rg->d_it_off.p; a.items = rg->d_items.p;
        a.nitems = rg->it_off.empty() ? 0 : rg->it_off.back();
        a.grp_off = rg->d_grp_off.p;
        a.ng4 = rg->ng4;
        a.images = nl->d_images.p; a.projs = nl->d_projs.p; a.nproj = (int)nl->geo.projs.size();
        a.coef = coef; a.coef_energy = nl->d_coef_energy.p; a.atom_coef_off = nl->d_atom_coef_off.p;
        a.ncoef = nl->geo.ncoef;
        a.max_col = std::max(1, rg->max_col);
    }
    for (int d = 0; d < 3; ++d) a.h[d] = g.h[d];
    const int nthr = NBW * a.tpb;
    QG_REQUIRE(nthr <= (NBW >= 4 ? 1024 : 512), "row tile too wide for one workgroup");
    QG_REQUIRE(!rg || rg->max_items <= nthr * kItemsPT, "a tile plane holds more projector items than staged");
    QG_REQUIRE(!rg || rg->ty * rg->ng4 + 1 <= nthr * kGroupsPT, "a tile plane holds more groups than staged");
    const size_t lds = rows_hpsi_lds<T>(g, sh.ty, sh.tz, NBW, NJS, a.max_col, a.nproj, nthr, a.ng4, PF);
    QG_REQUIRE(lds <= lds_limit(), "row tile does not fit in LDS");
    const int64_t nwg = (int64_t)a.ntile * a.nslab * ((nbands + NBW - 1) / NBW);
    QG_REQUIRE(nwg < (1LL << 31), "too many workgroups");
    if (rg)
        hipLaunchKernelGGL((rows_hpsi_kernel<T, NBW, NJS, PF, true>), dim3((unsigned)nwg), dim3(nthr), lds, s, a,
                           coef8<T>(g, lap_coef), V, psi, out);
    else
        hipLaunchKernelGGL((rows_hpsi_kernel<T, NBW, 4, 0, false>), dim3((unsigned)nwg), dim3(nthr), lds, s, a,
                           coef8<T>(g, lap_coef), V, psi, out);
    QG_HIP(hipGetLastError());
}

template <class T, int NJS, int PF>
static void rows_launch_hpsi(const qg_grid& g, const RowsShape& sh, const RowsGeo* rg, qg_nonlocal* nl, int nbands,
                             const T* psi, int64_t ld_in, T* out, int64_t ld_out, double lap_coef, int add_v,
                             int accumulate, const T* V, const double* coef, hipStream_t s) {
    switch (sh.nbw) {
        case 4:
            rows_launch_hpsi_t<T, 4, NJS, PF>(g, sh, rg, nl, nbands, psi, ld_in, out, ld_out, lap_coef, add_v,
                                               accumulate, V, coef, s);
            break;
        case 2:
            rows_launch_hpsi_t<T, 2, NJS, PF>(g, sh, rg, nl, nbands, psi, ld_in, out, ld_out, lap_coef, add_v,
                                               accumulate, V, coef, s);
            break;
        default:
            rows_launch_hpsi_t<T, 1, NJS, PF>(g, sh, rg, nl, nbands, psi, ld_in, out, ld_out, lap_coef, add_v,
                                               accumulate, V, coef, s);
    }
}

// The row path of one application.  Returns false (having launched nothing) when it does not
// serve this grid / projector set, so the caller runs the table kernels instead.
template <class T>
static bool rows_apply_t(qg_ham* h, int nbands, const T* psi, int64_t ld, T* hpsi, int64_t ld_out, hipStream_t s,
                         RowsScratch* ws) {
    const qg_grid& g = h->grid;
    qg_nonlocal* nl = (h->nl && h->nl->geo.ncoef > 0) ? h->nl : nullptr;
    const int dtype = sizeof(T) == 4 ? QG_F32 : QG_F64;
    if (!rows_eligible(g, dtype, nl)) return false;
    const RowsShape sh = rows_shape(g, dtype, nbands, &h->plan);
    const RowsGeo* rg = nullptr;    // apply kernel's geometry (tile rows ty)
    const RowsGeo* rgp = nullptr;   // projection kernel's geometry (tile rows typ)
    if (nl) {
        if (sh.typ == 0) return false;
        // with caller scratch nothing may be allocated: the geometry must exist already
        rg = rows_geometry(nl, sh.ty, sh.tzp, true, ws == nullptr);
        rgp = rows_geometry(nl, sh.typ, sh.tzp, false, ws == nullptr);
        QG_REQUIRE(rg && rgp, "the projector geometry of this launch shape is not prepared (qg_nonlocal_create and "
                              "qg_nonlocal_update prepare the default shapes)");
        if (!rg->ok || !rgp->ok) return false;
        const int nthr = rows_nthr(g, dtype, sh.ty, sh.nbw);
        if (rgp->max_items > kProjThreads * kItemsPT || rg->max_items > nthr * kItemsPT) return false;
        if (rg->ty * rg->ng4 + 1 > nthr * kGroupsPT) return false;
        if (rows_proj_lds<T>(*rgp, g, sh.nbp, (int)nl->geo.projs.size()) > lds_limit()) return false;
        if (rows_hpsi_lds<T>(g, sh.ty, sh.tz, sh.nbw, rg->njs, std::max(1, rg->max_col), (int)nl->geo.projs.size(),
                             nthr, rg->ng4, rg->pform) > lds_limit())
            return false;
    } else if (rows_hpsi_lds<T>(g, sh.ty, sh.tz, sh.nbw, 4, 1, 0, rows_nthr(g, dtype, sh.ty, sh.nbw), 0, 0) >
               lds_limit()) {
        return false;
    }
    const T* V = sizeof(T) == 4 ? (const T*)(const void*)h->V32.p : (const T*)(const void*)h->V64.p;
    double* part = nullptr;
    double* coef = nullptr;
    if (nl) {
        const size_t npart = (size_t)std::max(1, rgp->nslot) * rgp->njs * nbands;
        const size_t ncoef = (size_t)nbands * nl->geo.ncoef;
        if (ws) {
            QG_REQUIRE(npart + ncoef <= ws->capacity, "application exceeds the fixed scratch size");
            part = ws->base;
            coef = ws->base + npart;
        } else {
            h->rows_part.alloc(npart);
            h->rows_coef.alloc(ncoef);
            part = h->rows_part.p;
            coef = h->rows_coef.p;
        }
    }
    const bool timed = h->timing && ws == nullptr;
    if (timed) QG_HIP(hipEventRecord(h->tev[0], s));
    if (nl) {
        if (rgp->njs == 4) {
            if (rgp->pform == 2) rows_launch_project<T, 4, 2>(nl, *rgp, sh.nbp, nbands, psi, ld, part, coef, s);
            else if (rgp->pform == 1) rows_launch_project<T, 4, 1>(nl, *rgp, sh.nbp, nbands, psi, ld, part, coef, s);
            else rows_launch_project<T, 4, 0>(nl, *rgp, sh.nbp, nbands, psi, ld, part, coef, s);
        } else {
            rows_launch_project<T, 8, 0>(nl, *rgp, sh.nbp, nbands, psi, ld, part, coef, s);
        }
    }
    if (timed) QG_HIP(hipEventRecord(h->tev[1], s));
    if (!nl || rg->njs == 4) {
        if (nl && rg->pform == 2)
            rows_launch_hpsi<T, 4, 2>(g, sh, rg, nl, nbands, psi, ld, hpsi, ld_out, -0.5, 1, 0, V, coef, s);
        else if (nl && rg->pform == 1)
            rows_launch_hpsi<T, 4, 1>(g, sh, rg, nl, nbands, psi, ld, hpsi, ld_out, -0.5, 1, 0, V, coef, s);
        else
            rows_launch_hpsi<T, 4, 0>(g, sh, rg, nl, nbands, psi, ld, hpsi, ld_out, -0.5, 1, 0, V, coef, s);
    } else {
        rows_launch_hpsi<T, 8, 0>(g, sh, rg, nl, nbands, psi, ld, hpsi, ld_out, -0.5, 1, 0, V, coef, s);
    }
    if (timed) QG_HIP(hipEventRecord(h->tev[2], s));
    return true;
}

bool rows_hamiltonian_apply(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi,
                            int64_t ld_out, hipStream_t s, RowsScratch* ws) {
    if (h->plan.kernels == QG_KERNELS_TABLES) return false;
    if (h->timing) {
        while (h->tev.size() < 3) {
            hipEvent_t e;
            QG_HIP(hipEventCreate(&e));
            h->tev.push_back(e);
        }
    }
    bool done;
    if (dtype == QG_F32)
        done = rows_apply_t<float>(h, nbands, (const float*)psi, ld, (float*)hpsi, ld_out, s, ws);
    else
        done = rows_apply_t<double>(h, nbands, (const double*)psi, ld, (double*)hpsi, ld_out, s, ws);
    QG_REQUIRE(done || h->plan.kernels != QG_KERNELS_ROWS,
               "the plan asks for the row kernels, which do not serve this grid and projector set");
    if (done && h->timing) {
        h->tchunks = 1;
        h->timed_valid = true;
    }
    return done;
}

// coef_dev[b * ncoef + j] for a band block through the tile-centric projection; false when
// the row path does not serve this projector set.
bool rows_nonlocal_project(qg_nonlocal* nl, int dtype, int nbands, const void* psi, int64_t ld, double* coef,
                           DevBuf<double>& part_buf, hipStream_t s) {
    const qg_grid& g = nl->gcopy;
    if (nl->geo.ncoef == 0 || !rows_eligible(g, dtype, nl)) return false;
    const RowsShape sh = rows_shape(g, dtype, nbands);
    if (sh.typ == 0) return false;
    const RowsGeo* rg = rows_geometry(nl, sh.typ, sh.tzp, false);
    if (!rg->ok || rg->max_items > kProjThreads * kItemsPT) return false;
    const size_t lds = dtype == QG_F32 ? rows_proj_lds<float>(*rg, g, sh.nbp, (int)nl->geo.projs.size())
                                       : rows_proj_lds<double>(*rg, g, sh.nbp, (int)nl->geo.projs.size());
    if (lds > lds_limit()) return false;
    part_buf.alloc((size_t)std::max(1, rg->nslot) * rg->njs * nbands);
    auto go = [&](auto tag) {
        using T = decltype(tag);
        const T* p = (const T*)psi;
        if (rg->njs == 4) {
            if (rg->pform == 2) rows_launch_project<T, 4, 2>(nl, *rg, sh.nbp, nbands, p, ld, part_buf.p, coef, s);
            else if (rg->pform == 1) rows_launch_project<T, 4, 1>(nl, *rg, sh.nbp, nbands, p, ld, part_buf.p, coef, s);
            else rows_launch_project<T, 4, 0>(nl, *rg, sh.nbp, nbands, p, ld, part_buf.p, coef, s);
        } else {
            rows_launch_project<T, 8, 0>(nl, *rg, sh.nbp, nbands, p, ld, part_buf.p, coef, s);
        }
    };
    if (dtype == QG_F32) go(float{});
    else go(double{});
    return true;
}

// out = (accumulate ? out : 0) + coef Lap8 psi through the row stencil; false when the grid
// does not suit it.
bool rows_laplacian_apply(const qg_grid& g, int dtype, int nbands, const void* psi, int64_t ld, void* out,
                          int64_t ld_out, double coef, int accumulate, hipStream_t s) {
    if (!rows_eligible(g, dtype, nullptr)) return false;
    const RowsShape sh = rows_shape(g, dtype, nbands);
    if (dtype == QG_F32) {
        if (rows_hpsi_lds<float>(g, sh.ty, sh.tz, sh.nbw, 4, 1, 0, rows_nthr(g, dtype, sh.ty, sh.nbw), 0, 0) >
            lds_limit())
            return false;
        rows_launch_hpsi<float, 4, 0>(g, sh, nullptr, nullptr, nbands, (const float*)psi, ld, (float*)out, ld_out,
                                          coef, 0, accumulate, nullptr, nullptr, s);
    } else {
        if (rows_hpsi_lds<double>(g, sh.ty, sh.tz, sh.nbw, 4, 1, 0, rows_nthr(g, dtype, sh.ty, sh.nbw), 0, 0) >
            lds_limit())
            return false;
        rows_launch_hpsi<double, 4, 0>(g, sh, nullptr, nullptr, nbands, (const double*)psi, ld, (double*)out,
                                           ld_out, coef, 0, accumulate, nullptr, nullptr, s);
    }
    return true;
}

// Builds the row geometry of the default launch shapes (FP32 and FP64) of nl, so that
// applications find it ready (and applications on caller scratch need not allocate).
void rows_prepare(qg_nonlocal* nl) {
    if (nl->geo.ncoef == 0) return;
    for (int dtype : {QG_F32, QG_F64}) {
        if (!rows_eligible(nl->gcopy, dtype, nl)) continue;
        const RowsShape sh = rows_shape(nl->gcopy, dtype, 0);
        if (sh.typ == 0) continue;
        rows_geometry(nl, sh.ty, sh.tzp, true);
        rows_geometry(nl, sh.typ, sh.tzp, false);
    }
}

// ------------------------------------------------ allocation-free application ----
// Fixed device scratch of qg_hamiltonian_apply_noalloc: the FP64 projector partials and
// coefficients of one band chunk.  Every GPU the library runs on holds one copy of it (64 MiB)
// until the process ends; calls on one GPU share it, so calls that overlap on one GPU race.
constexpr size_t kScratchDoubles = size_t(8) << 20;
__device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
double* g_scratch_addr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device, looked up once per device (the array's
// name in host code is only a host-side stand-in for the symbol).
double* scratch_address() {
    int dev = 0;
    QG_HIP(hipGetDevice(&dev));
    QG_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        QG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        QG_REQUIRE(p != nullptr, "no device address for the scratch array");
        g_scratch_addr[dev] = static_cast<double*>(p);
    }
    return g_scratch_addr[dev];
}

// H psi through the row kernels with the fixed scratch: band chunks sized so that one chunk's
// partials and coefficients fit it, launched in stream order.  No device memory is allocated
// or freed; the projector geometry must have been prepared (qg_nonlocal_create/update do).
void rows_apply_noalloc(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi, int64_t ld_out,
                        hipStream_t s) {
    const qg_grid& g = h->grid;
    qg_nonlocal* nl = (h->nl && h->nl->geo.ncoef > 0) ? h->nl : nullptr;
    QG_REQUIRE(rows_eligible(g, dtype, nl),
               "the allocation-free H psi runs the row kernels only: x rows of whole 16-byte vectors (Nx a multiple "
               "of 4 in FP32, of 2 in FP64, 8 <= Nx <= 1024), Gaussian projectors with l <= 1, at most 8 "
               "coefficients per atom");
    RowsScratch ws{scratch_address(), kScratchDoubles};
    int chunk = nbands;
    if (nl) {
        const RowsShape sh = rows_shape(g, dtype, nbands, &h->plan);
        QG_REQUIRE(sh.typ > 0, "the projection tile does not fit this grid");
        const RowsGeo* rgp = rows_geometry(nl, sh.typ, sh.tzp, false, false);
        QG_REQUIRE(rgp && rgp->ok, "the projector geometry of this launch shape is not prepared (qg_nonlocal_create "
                                   "and qg_nonlocal_update prepare the default shapes)");
        const size_t per_band = (size_t)std::max(1, rgp->nslot) * rgp->njs + (size_t)nl->geo.ncoef;
        QG_REQUIRE(per_band <= kScratchDoubles,
                   "one band needs " + std::to_string(per_band) + " doubles of projector scratch, more than the fixed " +
                       std::to_string(kScratchDoubles));
        chunk = (int)std::min<size_t>((size_t)nbands, kScratchDoubles / per_band);
    }
    const size_t es = dtype == QG_F32 ? 4 : 8;
    for (int b0 = 0; b0 < nbands; b0 += chunk) {
        const int nb = std::min(chunk, nbands - b0);
        const char* pin = (const char*)psi + (size_t)b0 * ld * es;
        char* pout = (char*)hpsi + (size_t)b0 * ld_out * es;
        bool done;
        if (dtype == QG_F32)
            done = rows_apply_t<float>(h, nb, (const float*)pin, ld, (float*)pout, ld_out, s, &ws);
        else
            done = rows_apply_t<double>(h, nb, (const double*)pin, ld, (double*)pout, ld_out, s, &ws);
        QG_REQUIRE(done, "the row kernels do not serve this grid and projector set (tile does not fit in LDS)");
    }
}

}  // namespace qg
