// qg_hpsi.hip — the H·psi hot path on MI355X (gfx950): the fused 8th-order Laplacian +
// local potential + Kleinman-Bylander projector apply over a block of bands, the projector
// dot-product kernel, and their C-ABI entry points.
//
// Data layout in HBM: band-major blocks, x fastest (psi[b*ld + ix + Nx*(iy + Ny*iz)]).
// Projector values beta_j = p(r) Y_lm live in tables over the sphere points only (the
// reference's per-block projector and Ylm tables), flattened image by image in (z, y, x)
// order: 1.51 Ng points x 4 slots for the s+p workloads.
//
// Fused kernel (one pass over psi and H psi):
//   * a workgroup owns TY full x-rows, a slab of TZ z-planes and NB bands; it streams
//     through the slab in z with each point's 9 z-neighbours in a register queue and the
//     current plane (plus a 4-row y-halo and 4 wrapped x-columns) in LDS, double-buffered,
//     with the next plane's loads issued one plane ahead;
//   * the projector term of a plane is computed item-parallel: every sphere point of every
//     segment (image x tile row) reaching the plane is one work item; its table entry is
//     combined with E*coef (staged per workgroup) and added to a per-point LDS accumulator
//     with LDS atomics (spheres overlap), so no lane waits on another's sphere;
//   * workgroup ids are remapped so each XCD gets a contiguous range (y tile fastest), and
//     neighbouring tiles share that XCD's L2 for halo rows, V and projector tables.
// Projector coefficients come from a table-driven, atom-centric kernel.  A block of bands is
// processed in chunks of about 128 MB (projection, then fused apply), so the fused kernel
// re-reads each chunk from the 256 MB Infinity Cache rather than from HBM.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>
#include <type_traits>
#include "qg_internal.hpp"

namespace qg {

template <class T>
struct StencilCoef {
    T c0;
    T cx[4], cy[4], cz[4];
};

enum { MODE_LAP = 0, MODE_HAM = 1, MODE_NL = 2 };

struct FusedArgs {
    int nx, ny, nz;
    int ty, tz;
    int ntile_y, nslab;
    int nbands;
    int64_t ld_in, ld_out;
    int accumulate;
    // projector term (MODE_HAM with projectors, MODE_NL)
    int has_nl;
    const int* wg_atom_off;
    const int* wg_atom;
    const int64_t* item_off;
    const uint32_t* item_meta;
    const void* item_beta;     // T, njs values per item
    const double* coef;        // nbands x ncoef projections of this chunk
    const double* coef_energy;
    const int* atom_coef_off;
    int ncoef;
    int max_atom;
    int nlp;                   // z-planes per projector phase (accumulator ring depth, 1..4)
    // in-kernel projector values (every projector Gaussian): sphere segments per tile row and
    // the images / projectors they refer to
    int tile_major;            // workgroup order: tile fastest (no projector term) or band fastest
    int analytic;
    int max_img;
    int max_seg;               // most sphere segments of one (tile, slab)
    int nproj;
    const int* seg_off;
    const uint32_t* seg;
    const int* wg_img_off;
    const int* wg_img;
    const int* wg_img_atom;
    const DevImage* images;
    const DevProj* projs;
    double hx, hy, hz;
};

// Buffer-resource access: SGPR base + 32-bit byte offsets (VGPR per point, SGPR per plane).
using rsrc_t = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
template <class T>
__device__ __forceinline__ T bload(rsrc_t r, int voff, int soff) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    }
}
template <class T>
__device__ __forceinline__ void bstore(rsrc_t r, T v, int voff, int soff) {
    if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
    } else {
        using u2 = __attribute__((ext_vector_type(2))) uint32_t;
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, 0);
    }
}

// Map the hardware workgroup id (dispatched round-robin over the 8 XCDs) to a logical id
// such that each XCD receives one contiguous range of logical ids.
__device__ __forceinline__ int xcd_remap(int i, int n) {
    const int xcd = i & 7, slot = i >> 3;
    const int per = n >> 3, rem = n & 7;
    const int start = (xcd < rem) ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    return start + slot;
}

constexpr int kMaxThreads = 512;
constexpr int kHaloPT = 4;  // halo points per thread and plane (8*nx <= 4*threads)
constexpr int kItemPT = 4;  // projector work items per thread per round

// Fused kernel.  T = float/double, NB bands per workgroup, PPT points per thread, NJS table
// slots per projector item.  blockDim.x = threads (multiple of 64, <= 512).
// NB consecutive LDS values (16-byte aligned rows of the E*coef table) as 16-byte reads
template <class T, int NB>
__device__ __forceinline__ void lds_row(const T* __restrict__ src, T (&dst)[NB]) {
    constexpr int V = 16 / sizeof(T);
    if constexpr (NB % V == 0) {
        using Vec = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
        const Vec* s4 = reinterpret_cast<const Vec*>(src);
#pragma unroll
        for (int i = 0; i < NB / V; ++i) {
            const Vec v = s4[i];
            if constexpr (sizeof(T) == 4) {
                dst[4 * i] = v.x; dst[4 * i + 1] = v.y; dst[4 * i + 2] = v.z; dst[4 * i + 3] = v.w;
            } else {
                dst[2 * i] = v.x; dst[2 * i + 1] = v.y;
            }
        }
    } else {
#pragma unroll
        for (int b = 0; b < NB; ++b) dst[b] = src[b];
    }
}

template <class T, int NB, int PPT, int MODE, int NJS>
__device__ __forceinline__ void fused_body(const FusedArgs& a, const StencilCoef<T>& sc, const T* __restrict__ V,
                                           const T* __restrict__ psi, T* __restrict__ out, const int wg,
                                           unsigned char* smem_raw) {
    constexpr int S = sizeof(T);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int nthr = blockDim.x;
    const int W = nx + 8;                 // LDS row: 4 wrapped columns each side
    const int lds_band = (a.ty + 8) * W;
    const int tile_pts = a.ty * nx;
    // LDS carve-up
    unsigned char* lp = smem_raw;
    T* planes = reinterpret_cast<T*>(lp);
    if (MODE != MODE_NL) lp += 2 * (size_t)NB * lds_band * S;
    T* nl_l = reinterpret_cast<T*>(lp);       // per-point projector term, nlp x NB x tile_pts
    T* ec_l = nullptr;                        // E*coef per (local atom, slot, band)
    const int nlp = a.nlp;
    // in-kernel projector values exist in the projector pass only (the fused pass keeps its
    // registers for the stencil queue)
    constexpr bool kAnalyticMode = MODE == MODE_NL;
    const bool analytic = kAnalyticMode && a.has_nl && a.analytic;
    // in-kernel projector values: per image the fractional part of its centre in grid units
    // (3 x T) and integer part + shift, atom slot, projector range (6 ints); per projector
    // amplitude and 1/(2 sigma^2)
    // (2 x T), l and coefficient slot (2 ints); the slab's segment offsets and segments
    T* im_o = nullptr;
    T* pr_f = nullptr;
    int* im_i = nullptr;
    int* pr_i = nullptr;
    int* sg_off = nullptr;
    uint32_t* sg = nullptr;
    if (a.has_nl) {
        if (!analytic) lp += (size_t)nlp * NB * tile_pts * S;
        ec_l = reinterpret_cast<T*>(lp);
        lp += (size_t)a.max_atom * NJS * NB * S;
        if (analytic) {
            lp = smem_raw + (((size_t)(lp - smem_raw) + 7) & ~size_t(7));
            im_o = reinterpret_cast<T*>(lp);
            lp += (size_t)a.max_img * 3 * S;
            pr_f = reinterpret_cast<T*>(lp);
            lp += (size_t)a.nproj * 2 * S;
            im_i = reinterpret_cast<int*>(lp);
            lp += (size_t)a.max_img * 6 * sizeof(int);
            pr_i = reinterpret_cast<int*>(lp);
            lp += (size_t)a.nproj * 2 * sizeof(int);
            sg_off = reinterpret_cast<int*>(lp);
            lp += (size_t)(a.tz * a.ty + 1) * sizeof(int);
            sg = reinterpret_cast<uint32_t*>(lp);
        }
    }

    // logical order with a projector term: band chunk fastest, so the workgroups of one
    // (tile, slab) run back to back on one XCD and find V and the projector data in its L2
    const int nchunk = (a.nbands + NB - 1) / NB;
    int chunk, ytile, slab;
    if (a.tile_major) {
        // no projector term: neighbouring tiles and slabs of one band run back to back on one
        // XCD, so their shared halo rows and warm-up planes come from its L2
        ytile = wg % a.ntile_y;
        const int rest = wg / a.ntile_y;
        slab = rest % a.nslab;
        chunk = rest / a.nslab;
    } else {
        chunk = wg % nchunk;
        const int rest = wg / nchunk;
        ytile = rest % a.ntile_y;
        slab = rest / a.ntile_y;
    }
    const int y0 = ytile * a.ty;
    const int rows = min(a.ty, ny - y0);
    const int z0 = slab * a.tz;
    const int z1 = min(z0 + a.tz, nz);
    const int b0 = chunk * NB;
    const int tid = threadIdx.x;
    const int npts = rows * nx;
    const int plane = nx * ny;
    const uint32_t band_bytes = (uint32_t)((int64_t)plane * nz * S);

    rsrc_t rsrc_in[NB], rsrc_out[NB];
    bool bv[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        bv[b] = (b0 + b) < a.nbands;
        const int bb = bv[b] ? b0 + b : a.nbands - 1;
        rsrc_in[b] = make_rsrc(psi + (int64_t)bb * a.ld_in, band_bytes);
        rsrc_out[b] = make_rsrc(out + (int64_t)bb * a.ld_out, band_bytes);
    }
    const rsrc_t rsrc_v = make_rsrc(V, band_bytes);

    int vo[PPT];    // byte offset of the point inside a plane
    int lc[PPT];    // LDS index of the point (row iy+4, column ix+4), -1 when idle
    int ldup[PPT];  // LDS index of its wrapped copy in the x halo, or -1
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int p = tid + nthr * k;
        const bool ok = p < npts;
        const int iy = ok ? p / nx : 0;
        const int ix = ok ? p - iy * nx : 0;
        vo[k] = ok ? (y0 * nx + p) * S : 0;
        lc[k] = ok ? (iy + 4) * W + ix + 4 : -1;
        ldup[k] = (ok && ix < 4) ? lc[k] + nx : ((ok && ix >= nx - 4) ? lc[k] - nx : -1);
    }
    // halo elements of this thread: 8 rows x nx (4 above, 4 below the tile), wrapped in y
    int ho[kHaloPT], hl[kHaloPT];
#pragma unroll
    for (int j = 0; j < kHaloPT; ++j) {
        const int e = tid + nthr * j;
        const bool ok = (MODE != MODE_NL) && e < 8 * nx;
        const int r8 = ok ? e / nx : 0;
        const int col = ok ? e - r8 * nx : 0;
        int gy = (r8 < 4) ? y0 - 4 + r8 : y0 + rows + (r8 - 4);
        gy = ((gy % ny) + ny) % ny;
        const int ly = (r8 < 4) ? r8 : rows + 4 + (r8 - 4);
        ho[j] = ok ? (gy * nx + col) * S : 0;
        hl[j] = ok ? ly * W + col + 4 : -1;
    }
    auto zoff = [&](int z) { int zz = z % nz; zz = zz < 0 ? zz + nz : zz; return zz * plane * S; };

    // ---- projector staging for the workgroup: E*coef per (atom slot, j, band) ----
    const int wgi = ytile * a.nslab + slab;
    if (a.has_nl) {
        const int a0 = a.wg_atom_off[wgi];
        const int natom = a.wg_atom_off[wgi + 1] - a0;
        for (int i = tid; i < natom * NJS * NB; i += nthr) {
            const int as = i / (NJS * NB), j = (i / NB) % NJS, b = i % NB;
            const int at = a.wg_atom[a0 + as];
            const int cb = a.atom_coef_off[at];
            const int cnt = a.atom_coef_off[at + 1] - cb;
            T v = T(0);
            if (j < cnt && (b0 + b) < a.nbands)
                v = (T)(a.coef_energy[cb + j] * a.coef[(int64_t)(b0 + b) * a.ncoef + cb + j]);
            ec_l[i] = v;
        }
        if (analytic) {
            const int i0 = a.wg_img_off[wgi];
            const int nimg = a.wg_img_off[wgi + 1] - i0;
            const double hh[3] = {a.hx, a.hy, a.hz};
            for (int li = tid; li < nimg; li += nthr) {
                const DevImage& im = a.images[a.wg_img[i0 + li]];
                // centre in grid units u = R / h = n + f (n nearest integer): the displacement
                // h ((i - shift - n) - f) is then exact to the last bit of f, so a nucleus on a
                // grid point gives r^2 ~ 1e-30 there (Y_1m = 0, as in the host tables) in FP32 too
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double u = im.R[d] / hh[d];
                    const double nn = rint(u);
                    im_o[3 * li + d] = (T)(u - nn);
                    im_i[6 * li + d] = im.shift[d] + (int)nn;
                }
                im_i[6 * li + 3] = a.wg_img_atom[i0 + li];
                im_i[6 * li + 4] = im.proj_begin;
                im_i[6 * li + 5] = im.proj_end;
            }
            for (int pj = tid; pj < a.nproj; pj += nthr) {
                const DevProj& d = a.projs[pj];
                pr_f[2 * pj] = (T)d.amp;
                pr_f[2 * pj + 1] = (T)d.inv_2s2;
                pr_i[2 * pj] = d.l;
                pr_i[2 * pj + 1] = d.coef_rel;
            }
            // segments of the slab's rows: [sg_off[(z - z0) * ty + row], ... + 1)
            const size_t k0 = ((size_t)ytile * nz + z0) * a.ty;
            const int nrow = (z1 - z0) * a.ty;
            const int sbase = a.seg_off[k0];
            for (int i = tid; i <= nrow; i += nthr) sg_off[i] = a.seg_off[k0 + i] - sbase;
            const int nseg = a.seg_off[k0 + nrow] - sbase;
            for (int i = tid; i < nseg; i += nthr) sg[i] = a.seg[sbase + i];
        } else {
            for (int i = tid; i < nlp * NB * tile_pts; i += nthr) nl_l[i] = T(0);
        }
    }

    T q[NB][PPT][9];
    T qn[NB][PPT];
    T hn[NB][kHaloPT];
    T vn[PPT];
    if constexpr (MODE != MODE_NL) {
#pragma unroll
        for (int d = 0; d < 8; ++d) {
            const int so = zoff(z0 - 4 + d);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int k = 0; k < PPT; ++k) q[b][k][d] = bload<T>(rsrc_in[b], vo[k], so);
        }
    }
    auto prefetch = [&](int z) {   // stencil inputs of plane z: psi(z+4), halo(z), V(z)
        if constexpr (MODE != MODE_NL) {
            const int so = zoff(z + 4), sh = zoff(z);
#pragma unroll
            for (int b = 0; b < NB; ++b) {
#pragma unroll
                for (int k = 0; k < PPT; ++k) qn[b][k] = bload<T>(rsrc_in[b], vo[k], so);
#pragma unroll
                for (int j = 0; j < kHaloPT; ++j) hn[b][j] = (hl[j] >= 0) ? bload<T>(rsrc_in[b], ho[j], sh) : T(0);
            }
        }
        if constexpr (MODE == MODE_HAM) {
            const int sv = zoff(z);
#pragma unroll
            for (int k = 0; k < PPT; ++k) vn[k] = bload<T>(rsrc_v, vo[k], sv);
        }
    };
    prefetch(z0);
    const T* ibeta = reinterpret_cast<const T*>(a.item_beta);

    for (int z = z0; z < z1; ++z) {
        const int buf = (z - z0) & 1;
        const int sz = z * plane * S;
        T vcur[PPT];
        // commit the prefetched plane into registers / LDS
        if constexpr (MODE != MODE_NL) {
            T* L = planes + buf * NB * lds_band;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    q[b][k][8] = qn[b][k];
                    if (lc[k] >= 0) L[b * lds_band + lc[k]] = q[b][k][4];
                    if (ldup[k] >= 0) L[b * lds_band + ldup[k]] = q[b][k][4];
                }
#pragma unroll
                for (int j = 0; j < kHaloPT; ++j)
                    if (hl[j] >= 0) L[b * lds_band + hl[j]] = hn[b][j];
            }
        }
#pragma unroll
        for (int k = 0; k < PPT; ++k) vcur[k] = (MODE == MODE_HAM) ? vn[k] : T(0);
        // issue the loads of the next plane
        if (z + 1 < z1) prefetch(z + 1);
        __syncthreads();

        // projector term: every (image, sphere point) reaching the next nlp planes is one work
        // item, accumulated into the ring slot of its plane
        if (MODE != MODE_LAP && a.has_nl && !analytic && (z - z0) % nlp == 0) {
            const size_t tp = (size_t)ytile * nz + z;
            const int zend = min(z + nlp, z1);
            const int64_t io0 = a.item_off[tp];
            const int64_t io_end = a.item_off[(size_t)ytile * nz + zend];
            const int nitem = (int)(io_end - io0);
            int rb[3];   // item index (relative to io0) where planes z+1, z+2, z+3 start
#pragma unroll
            for (int k = 0; k < 3; ++k) rb[k] = (z + k + 1 < zend) ? (int)(a.item_off[tp + k + 1] - io0) : nitem;
            for (int base = 0; base < nitem; base += nthr * kItemPT) {
                uint32_t meta[kItemPT];
                T bj[kItemPT][NJS];
#pragma unroll
                for (int u = 0; u < kItemPT; ++u) {
                    const int it = base + tid + nthr * u;
                    meta[u] = 0xffffffffu;
                    if (it < nitem) {
                        meta[u] = a.item_meta[io0 + it];
                        const T* bt = ibeta + (io0 + it) * NJS;
#pragma unroll
                        for (int j = 0; j < NJS; ++j) bj[u][j] = bt[j];
                    }
                }
#pragma unroll
                for (int u = 0; u < kItemPT; ++u) {
                    if (meta[u] == 0xffffffffu) continue;
                    const int pt = (int)(meta[u] & 0xffffu);
                    const T* ec = ec_l + (int)(meta[u] >> 16) * NJS * NB;
                    const int it = base + tid + nthr * u;
                    const int pr = (it >= rb[0]) + (it >= rb[1]) + (it >= rb[2]);
                    T* acc_base = nl_l + (size_t)pr * NB * tile_pts;
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        T acc = T(0);
#pragma unroll
                        for (int j = 0; j < NJS; ++j) acc += ec[j * NB + b] * bj[u][j];
                        atomicAdd(&acc_base[b * tile_pts + pt], acc);
                    }
                }
            }
            __syncthreads();
        }

        T res[NB][PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if constexpr (MODE == MODE_HAM) {
                    res[b][k] = vcur[k] * q[b][k][4];
                } else {
                    res[b][k] = (a.accumulate && lc[k] >= 0) ? bload<T>(rsrc_out[b], vo[k], sz) : T(0);
                }
                if (MODE != MODE_LAP && a.has_nl && !analytic && lc[k] >= 0) {
                    T* slot = &nl_l[((size_t)((z - z0) % nlp) * NB + b) * tile_pts + tid + nthr * k];
                    res[b][k] += *slot;
                    *slot = T(0);
                }
            }
        if (kAnalyticMode && analytic) {
            // beta_j = p(r) Y_lm of every sphere this point lies in, evaluated here in the
            // block's precision (displacement from the integer part of the image centre, so
            // every term is a few grid spacings at most), and acc_b = sum_j E_j c_jb beta_j
            const T hx = (T)a.hx, hy = (T)a.hy, hz = (T)a.hz;
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                if (lc[k] < 0) continue;
                const int p = tid + nthr * k;
                const int ly = p / nx, ix = p - ly * nx;
                const int iy = y0 + ly;
                const int rk = (z - z0) * a.ty + ly;
                const int s0 = sg_off[rk], s1 = sg_off[rk + 1];
                T acc[NB];
#pragma unroll
                for (int b = 0; b < NB; ++b) acc[b] = T(0);
                // the row's segments that contain this point, as a bit mask over up to 64 of
                // them at a time: the evaluation loop then runs once per membership (about 1.5
                // per point) instead of once per segment of the row
                for (int c0 = s0; c0 < s1; c0 += 64) {
                  const int c1 = min(s1, c0 + 64);
                  uint64_t mask = 0;
                  for (int si = c0; si < c1; ++si) {
                      const uint32_t sv = sg[si];
                      const int x0 = (int)(sv & 0x7ffu), x1 = (int)((sv >> 11) & 0x7ffu);
                      mask |= (uint64_t)(ix >= x0 && ix < x1) << (si - c0);
                  }
                  while (mask) {
                    const int si = c0 + __builtin_ctzll(mask);
                    mask &= mask - 1;
                    const int li = (int)(sg[si] >> 22);
                    const int* ii = im_i + 6 * li;
                    const T* io = im_o + 3 * li;
                    const T xc = hx * ((T)(ix - ii[0]) - io[0]);
                    const T yc = hy * ((T)(iy - ii[1]) - io[1]);
                    const T zc = hz * ((T)(z - ii[2]) - io[2]);
                    const T rr = xc * xc + (yc * yc + zc * zc);
                    const T* ec = ec_l + ii[3] * NJS * NB;
                    T yv[3] = {T(0), T(0), T(0)};
                    if (!(T(1.0e-14) > rr)) {
                        // sqrt(3/4pi) / r; FP32 through the hardware reciprocal square root
                        // (about 1 ulp), FP64 with a correctly rounded division
                        T ir;
                        if constexpr (sizeof(T) == 4) ir = T(0.48860251190291992) * __frsqrt_rn(rr);
                        else ir = T(0.48860251190291992) / sqrt(rr);
                        yv[0] = yc * ir;
                        yv[1] = zc * ir;
                        yv[2] = xc * ir;
                    }
                    for (int pj = ii[4]; pj < ii[5]; ++pj) {
                        const T rad = pr_f[2 * pj] * exp(-(rr * pr_f[2 * pj + 1]));
                        const int cr = pr_i[2 * pj + 1];
                        if (pr_i[2 * pj] == 0) {
                            const T bv = rad * T(0.28209479177387814);   // 1/sqrt(4 pi)
                            T e0[NB];
                            lds_row<T, NB>(ec + cr * NB, e0);
#pragma unroll
                            for (int b = 0; b < NB; ++b) acc[b] += e0[b] * bv;
                        } else {
                            T e0[NB], e1[NB], e2[NB];
                            lds_row<T, NB>(ec + cr * NB, e0);
                            lds_row<T, NB>(ec + (cr + 1) * NB, e1);
                            lds_row<T, NB>(ec + (cr + 2) * NB, e2);
#pragma unroll
                            for (int b = 0; b < NB; ++b)
                                acc[b] += rad * (e0[b] * yv[0] + e1[b] * yv[1] + e2[b] * yv[2]);
                        }
                    }
                  }
                }
#pragma unroll
                for (int b = 0; b < NB; ++b) res[b][k] += acc[b];
            }
        }
        if constexpr (MODE != MODE_NL) {
            const T* L = planes + buf * NB * lds_band;
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                if (lc[k] < 0) continue;
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const T* Lb = L + b * lds_band + lc[k];
                    T d = sc.c0 * q[b][k][4];
                    d += sc.cx[3] * (Lb[4] + Lb[-4]) + sc.cx[2] * (Lb[3] + Lb[-3]) +
                         sc.cx[1] * (Lb[2] + Lb[-2]) + sc.cx[0] * (Lb[1] + Lb[-1]);
                    d += sc.cy[3] * (Lb[4 * W] + Lb[-4 * W]) + sc.cy[2] * (Lb[3 * W] + Lb[-3 * W]) +
                         sc.cy[1] * (Lb[2 * W] + Lb[-2 * W]) + sc.cy[0] * (Lb[W] + Lb[-W]);
                    d += sc.cz[3] * (q[b][k][8] + q[b][k][0]) + sc.cz[2] * (q[b][k][7] + q[b][k][1]) +
                         sc.cz[1] * (q[b][k][6] + q[b][k][2]) + sc.cz[0] * (q[b][k][5] + q[b][k][3]);
                    res[b][k] += d;
                }
            }
        }
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int k = 0; k < PPT; ++k)
                if (lc[k] >= 0 && bv[b]) bstore<T>(rsrc_out[b], res[b][k], vo[k], sz);
        if constexpr (MODE != MODE_NL) {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int k = 0; k < PPT; ++k)
#pragma unroll
                    for (int d = 0; d < 8; ++d) q[b][k][d] = q[b][k][d + 1];
        }
    }
}

template <class T, int NB, int PPT, int MODE, int NJS>
__global__ __launch_bounds__(kMaxThreads) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 3 : 4)))
void fused_hpsi_kernel(FusedArgs a, StencilCoef<T> sc, const T* __restrict__ V, const T* __restrict__ psi,
                       T* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    fused_body<T, NB, PPT, MODE, NJS>(a, sc, V, psi, out, xcd_remap(blockIdx.x, gridDim.x), smem_raw);
}

// One cell of a batched application (independent cells of different sizes in one launch).
template <class T>
struct BatchCell {
    FusedArgs a;
    StencilCoef<T> sc;
    const T* V;
    const T* psi;
    T* out;
};

// Persistent form over a batch of cells: the workgroups of all cells form one logical range
// (cell c owns [prefix[c], prefix[c+1])); each hardware workgroup walks it with a stride of
// the grid size, starting from its XCD-contiguous logical id.
template <class T, int NB, int PPT, int MODE, int NJS>
__global__ __launch_bounds__(kMaxThreads) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 3 : 4)))
void fused_batch_kernel(const BatchCell<T>* __restrict__ cells, const int* __restrict__ prefix, int ncells) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int total = prefix[ncells];
    for (int w = xcd_remap(blockIdx.x, gridDim.x); w < total; w += gridDim.x) {
        int lo = 0, hi = ncells - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= w) lo = mid;
            else hi = mid - 1;
        }
        const BatchCell<T>& c = cells[lo];
        fused_body<T, NB, PPT, MODE, NJS>(c.a, c.sc, c.V, c.psi, c.out, w - prefix[lo], smem_raw);
        __syncthreads();   // LDS is reused by the next item
    }
}

// ---------------------------------------------------------------------------------------
// Projector coefficients: coef[b*ncoef + cbase(a) + j] = dV * sum_i beta_j(i) psi_b(i).
// One workgroup per (atom, chunk of NBA bands), streaming the atom's table entries.
struct ProjArgs {
    int natoms;
    int nbands;
    int64_t ld;
    double dv;
    const int* atom_img_off;
    const int* atom_coef_off;
    const int64_t* img_tbl_off;
    const int* gidx;
    const void* beta;
    double* coef;
    int ncoef;
};

constexpr int kProjBlock = 256;

template <class T, int NBA, int NJS>
__device__ __forceinline__ void project_body(const ProjArgs& a, const T* __restrict__ psi, const int item,
                                             double (*red)[NJS * NBA]) {
    const int atom = item % a.natoms;
    const int chunk = item / a.natoms;
    const int b0 = chunk * NBA;
    const int tid = threadIdx.x;
    const T* beta = reinterpret_cast<const T*>(a.beta);
    const T* src[NBA];
#pragma unroll
    for (int b = 0; b < NBA; ++b) src[b] = psi + (int64_t)min(b0 + b, a.nbands - 1) * a.ld;
    T acc[NJS][NBA];
#pragma unroll
    for (int j = 0; j < NJS; ++j)
#pragma unroll
        for (int b = 0; b < NBA; ++b) acc[j][b] = T(0);
    const int64_t t0 = a.img_tbl_off[a.atom_img_off[atom]];
    const int64_t t1 = a.img_tbl_off[a.atom_img_off[atom + 1]];
    constexpr int U = 4;   // independent table points in flight per thread
    for (int64_t base = t0; base < t1; base += (int64_t)kProjBlock * U) {
        int g[U];
        bool ok[U];
        T bj[U][NJS];
        T v[U][NBA];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = base + tid + (int64_t)kProjBlock * u;
            ok[u] = i < t1;
            g[u] = ok[u] ? a.gidx[i] : 0;
#pragma unroll
            for (int j = 0; j < NJS; ++j) bj[u][j] = ok[u] ? beta[i * NJS + j] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int b = 0; b < NBA; ++b) v[u][b] = ok[u] ? src[b][g[u]] : T(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int b = 0; b < NBA; ++b)
#pragma unroll
                for (int j = 0; j < NJS; ++j) acc[j][b] += bj[u][j] * v[u][b];
    }
    const int wave = tid >> 6;
#pragma unroll
    for (int j = 0; j < NJS; ++j)
#pragma unroll
        for (int b = 0; b < NBA; ++b) {
            double v = (double)acc[j][b];
#pragma unroll
            for (int s = 32; s >= 1; s >>= 1) v += __shfl_xor(v, s, 64);
            if ((tid & 63) == 0) red[wave][j * NBA + b] = v;
        }
    __syncthreads();
    if (tid < NJS * NBA) {
        const int j = tid / NBA, b = tid % NBA;
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < kProjBlock / 64; ++w) sum += red[w][tid];
        const int cbase = a.atom_coef_off[atom];
        const int count = a.atom_coef_off[atom + 1] - cbase;
        if (b0 + b < a.nbands && j < count) a.coef[(int64_t)(b0 + b) * a.ncoef + cbase + j] = sum * a.dv;
    }
}

template <class T, int NBA, int NJS>
__global__ __launch_bounds__(kProjBlock) void project_kernel(ProjArgs a, const T* __restrict__ psi) {
    __shared__ double red[kProjBlock / 64][NJS * NBA];
    // consecutive items (neighbouring atoms of one band chunk, whose spheres share grid rows)
    // on one XCD, so the shared rows come from that XCD's L2
    project_body<T, NBA, NJS>(a, psi, xcd_remap(blockIdx.x, gridDim.x), red);
}

template <class T>
struct ProjCell {
    ProjArgs a;
    const T* psi;
};

// Persistent projection over a batch of cells: items (cell, atom, band chunk) in one range.
template <class T, int NBA, int NJS>
__global__ __launch_bounds__(kProjBlock) void project_batch_kernel(const ProjCell<T>* __restrict__ cells,
                                                                   const int* __restrict__ prefix, int ncells) {
    __shared__ double red[kProjBlock / 64][NJS * NBA];
    const int total = prefix[ncells];
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        int lo = 0, hi = ncells - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= w) lo = mid;
            else hi = mid - 1;
        }
        project_body<T, NBA, NJS>(cells[lo].a, cells[lo].psi, w - prefix[lo], red);
        __syncthreads();   // red is reused by the next item
    }
}

// ------------------------------------------------------------------ host launchers ----
template <class T>
static StencilCoef<T> make_coef(const qg_grid& g, double coef) {
    const double w[4] = {8064.0, -1008.0, 128.0, -9.0};
    StencilCoef<T> s;
    const double c0 = -14350.0 / 5040.0 * coef *
                      (1.0 / (g.h[0] * g.h[0]) + 1.0 / (g.h[1] * g.h[1]) + 1.0 / (g.h[2] * g.h[2]));
    s.c0 = (T)c0;
    for (int m = 0; m < 4; ++m) {
        s.cx[m] = (T)(w[m] * coef / (5040.0 * g.h[0] * g.h[0]));
        s.cy[m] = (T)(w[m] * coef / (5040.0 * g.h[1] * g.h[1]));
        s.cz[m] = (T)(w[m] * coef / (5040.0 * g.h[2] * g.h[2]));
    }
    return s;
}

struct Plan {
    int ty, tz, ppt, nb, nthr;
    int nlp = 1;
};

static int round64(int v) { return (v + 63) / 64 * 64; }


// Tile shapes chosen by the autotuner (qg_autotune_hamiltonian), per grid shape and dtype;
// make_plan uses them for the stencil modes unless QG_PLAN_* overrides are set.
struct TunedShape {
    int ty, ppt;
};
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int>, TunedShape> g_tuned;
static thread_local const TunedShape* t_trial = nullptr;   // shape under trial by the autotuner

static bool lookup_tuned(const qg_grid& g, int dtype, TunedShape& out) {
    if (t_trial) {
        out = *t_trial;
        return true;
    }
    if (std::getenv("QG_PLAN_TY") || std::getenv("QG_PLAN_PPT")) return false;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(std::make_tuple(g.n[0], g.n[1], g.n[2], dtype));
    if (it == g_tuned.end()) return false;
    out = it->second;
    return true;
}

// Tile plan: a tile is TY full x-rows; PPT points per thread; threads = ceil(TY*Nx/PPT)
// rounded up to whole wavefronts (<= 512); slabs of TZ planes, deep enough to amortise the
// 8-plane z warm-up, shallow enough to give >= ~2048 workgroups.  QG_PLAN_NB / _TY / _PPT /
// _TZ override the choice for tuning sweeps.
static Plan make_plan(const qg_grid& g, int dtype, int nbands, int mode) {
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    QG_REQUIRE(nx <= 128, "the fused kernel supports Nx <= 128");
    Plan p{};
    if (mode == MODE_NL) {
        // projector-only pass: no stencil state, so many bands share each projector item
        const int want = env_int("QG_PLAN_NBNL", dtype == QG_F64 ? 4 : 8);
        p.nb = 1;
        for (int c : {16, 8, 4, 2})
            if (c <= want && c <= nbands) { p.nb = c; break; }
        // measured at 80^3 FP32: 4-row tiles, 2 points per thread, 2-plane slabs, 8 bands
        p.ty = std::max(1, std::min(ny, env_int("QG_PLAN_TYNL", std::max(1, 320 / nx))));
        p.ppt = env_int("QG_PLAN_PPTNL", 2) == 4 ? 4 : 2;
        p.nthr = round64((p.ty * nx + p.ppt - 1) / p.ppt);
        while (p.nthr > kMaxThreads) {
            p.ppt = (p.ppt == 2) ? 4 : (p.ppt == 4) ? 5 : 8;
            p.nthr = round64((p.ty * nx + p.ppt - 1) / p.ppt);
            QG_REQUIRE(p.ppt <= 8 && (p.ppt < 8 || p.nthr <= kMaxThreads), "tile too large");
        }
        // no z warm-up in this pass: thin slabs give many workgroups
        p.tz = std::max(1, std::min(nz, env_int("QG_PLAN_TZNL", 2)));
        return p;
    }
    p.nb = env_int("QG_PLAN_NB", 1);
    p.nb = (p.nb == 2 && nbands >= 2) ? 2 : 1;
    // measured on MI355X at 80^3: narrow tiles (480 points FP32, 960 FP64) with 2 points per
    // thread keep more workgroups per CU in flight through the projector phases
    p.ty = std::max(1, std::min(ny, env_int("QG_PLAN_TY", 480 / nx)));
    p.ppt = env_int("QG_PLAN_PPT", 2);
    TunedShape ts;
    if (lookup_tuned(g, dtype, ts)) {
        p.ty = std::max(1, std::min(ny, ts.ty));
        p.ppt = ts.ppt;
    }
    QG_REQUIRE(p.ppt == 2 || p.ppt == 4 || p.ppt == 5 || p.ppt == 8, "QG_PLAN_PPT must be 2, 4, 5 or 8");
    p.nthr = round64((p.ty * nx + p.ppt - 1) / p.ppt);
    while (p.nthr > kMaxThreads) {
        p.ppt = (p.ppt == 2) ? 4 : (p.ppt == 4) ? 5 : 8;
        p.nthr = round64((p.ty * nx + p.ppt - 1) / p.ppt);
        QG_REQUIRE(p.ppt <= 8 && (p.ppt < 8 || p.nthr <= kMaxThreads), "tile too large");
    }
    // enough threads to stage the 8 halo rows (kHaloPT points each)
    p.nthr = std::max(p.nthr, round64((8 * nx + kHaloPT - 1) / kHaloPT));
    QG_REQUIRE(8 * nx <= kHaloPT * p.nthr && p.nthr <= kMaxThreads, "halo does not fit the thread block");
    const int ntile = (ny + p.ty - 1) / p.ty;
    const int nchunk = (nbands + p.nb - 1) / p.nb;
    int tz = nz;
    // slabs of 10-16 planes: short enough for many workgroups in flight, long enough for the
    // 8-plane z warm-up
    while (tz > 16 && (int64_t)ntile * nchunk * ((nz + tz - 1) / tz) < 65536) tz = (tz + 1) / 2;
    p.tz = std::max(1, std::min(nz, env_int("QG_PLAN_TZ", tz)));
    p.nlp = std::max(1, std::min(4, env_int("QG_PLAN_NLP", 4)));
    return p;
}

// Upload the per-plan staging lists of the projector term (cached per (ty, tz)).
static qg_nonlocal::Tiles& prepare_tile_nl(qg_nonlocal* nl, const Plan& p, bool segments) {
    qg_nonlocal::Tiles& c = nl->tiles[segments ? 1 : 0];
    if (c.ty == p.ty && c.tz == p.tz) return c;
    TileNL& t = c.tile;
    build_tile_nl(nl->gcopy, nl->geo, p.ty, p.tz, segments, t);
    auto up = [](auto& buf, auto& vec) {
        if (vec.empty()) vec.push_back({});
        buf.upload(vec.data(), vec.size());
    };
    up(c.d_wg_atom_off, t.wg_atom_off);
    up(c.d_wg_atom, t.wg_atom);
    up(c.d_wg_img_off, t.wg_img_off);
    up(c.d_wg_img, t.wg_img);
    up(c.d_wg_img_atom, t.wg_img_atom);
    up(c.d_seg_off, t.seg_off);
    up(c.d_seg, t.seg);
    up(c.d_item_off, t.item_off);
    up(c.d_item_meta, t.item_meta);
    up(c.d_item_beta64, t.item_beta);
    std::vector<float> b32(t.item_beta.begin(), t.item_beta.end());
    up(c.d_item_beta32, b32);
    // the host copies are not needed once uploaded
    std::vector<uint32_t>().swap(t.seg);
    std::vector<uint32_t>().swap(t.item_meta);
    std::vector<double>().swap(t.item_beta);
    c.ty = p.ty;
    c.tz = p.tz;
    return c;
}

// Projector form of a launch: in-kernel values for the FP32 projector pass when the handle
// allows them (QG_NL_ANALYTIC=2: for the FP64 projector pass too), per-point tables otherwise;
// the fused single pass always reads tables.
static bool segments_form(const qg_nonlocal* nl, int dtype, int mode) {
    if (!nl->analytic || mode != MODE_NL) return false;
    return dtype == QG_F32 || nl->analytic_level >= 2;
}

template <class T, int NB, int MODE, int NJS>
static void launch_fused_ppt(const Plan& p, dim3 grid, size_t lds, hipStream_t s, const FusedArgs& a,
                             const StencilCoef<T>& sc, const T* V, const T* psi, T* out) {
    switch (p.ppt) {
#define QG_CASE(PP)                                                                                 \
    case PP:                                                                                        \
        hipLaunchKernelGGL((fused_hpsi_kernel<T, NB, PP, MODE, NJS>), grid, dim3(p.nthr), lds, s, a, sc, V, \
                           psi, out);                                                               \
        break;
        QG_CASE(2) QG_CASE(4) QG_CASE(5) QG_CASE(8)
#undef QG_CASE
        default: throw Error{QG_ERR_ARG, "unsupported points-per-thread"};
    }
    QG_HIP(hipGetLastError());
}

template <class T>
struct FusedSetup {
    Plan p;
    FusedArgs a;
    StencilCoef<T> sc;
    size_t lds;
    int njs;
    int64_t nwg;
};

// Plan, kernel arguments and LDS size of one fused launch (also used per cell by the batch).
template <class T, int MODE>
static FusedSetup<T> setup_fused(const qg_grid& g, int nbands, int64_t ld_in, int64_t ld_out, double coef,
                                 int accumulate, qg_nonlocal* nl, const double* coefs) {
    const int dtype = sizeof(T) == 4 ? QG_F32 : QG_F64;
    Plan p = make_plan(g, dtype, nbands, MODE);
    QG_REQUIRE(g.ng() * (int64_t)sizeof(T) < (1LL << 31), "grid too large for 32-bit buffer offsets");
    FusedArgs a{};
    a.nx = g.n[0]; a.ny = g.n[1]; a.nz = g.n[2];
    a.ty = p.ty; a.tz = p.tz;
    a.ntile_y = (g.n[1] + p.ty - 1) / p.ty;
    a.nslab = (g.n[2] + p.tz - 1) / p.tz;
    a.nbands = nbands;
    a.ld_in = ld_in; a.ld_out = ld_out;
    a.accumulate = accumulate;
    a.has_nl = 0;
    a.nlp = p.nlp;
    size_t lds = (MODE == MODE_NL) ? 0 : 2 * (size_t)p.nb * (p.ty + 8) * (g.n[0] + 8) * sizeof(T);
    int njs = 4;
    if (nl && nl->geo.ncoef > 0 && MODE != MODE_LAP) {
        bool seg = segments_form(nl, dtype, MODE);
        if (seg) {
            // the in-kernel form stages the slab's images and segments; tables when they do not
            // fit in LDS or the segment encoding
            const TileNL& t = prepare_tile_nl(nl, p, true).tile;
            const size_t need = ((lds + (size_t)std::max(1, t.max_atom) * nl->geo.njs * p.nb * sizeof(T) + 7) & ~size_t(7)) +
                                (size_t)std::max(1, t.max_img) * (3 * sizeof(T) + 6 * sizeof(int)) +
                                nl->geo.projs.size() * (2 * sizeof(T) + 2 * sizeof(int)) +
                                (size_t)(p.tz * p.ty + 1) * sizeof(int) + (size_t)std::max(1, t.max_seg) * sizeof(uint32_t);
            if (need > 65536 || t.seg_off.empty()) seg = false;
        }
        qg_nonlocal::Tiles& tc = prepare_tile_nl(nl, p, seg);
        const TileNL& t = tc.tile;
        njs = nl->geo.njs;
        a.has_nl = 1;
        a.wg_atom_off = tc.d_wg_atom_off.p;
        a.wg_atom = tc.d_wg_atom.p;
        a.item_off = tc.d_item_off.p;
        a.item_meta = tc.d_item_meta.p;
        a.item_beta = (dtype == QG_F32) ? (const void*)tc.d_item_beta32.p : (const void*)tc.d_item_beta64.p;
        a.coef = coefs;
        a.coef_energy = nl->d_coef_energy.p;
        a.atom_coef_off = nl->d_atom_coef_off.p;
        a.ncoef = nl->geo.ncoef;
        a.max_atom = std::max(1, t.max_atom);
        const size_t ec = (size_t)a.max_atom * njs * p.nb * sizeof(T);
        if (seg) {
            a.analytic = 1;
            a.max_img = std::max(1, t.max_img);
            a.seg_off = tc.d_seg_off.p;
            a.seg = tc.d_seg.p;
            a.wg_img_off = tc.d_wg_img_off.p;
            a.wg_img = tc.d_wg_img.p;
            a.wg_img_atom = tc.d_wg_img_atom.p;
            a.images = nl->d_images.p;
            a.projs = nl->d_projs.p;
            a.hx = g.h[0]; a.hy = g.h[1]; a.hz = g.h[2];
            a.nproj = (int)nl->geo.projs.size();
            a.max_seg = std::max(1, t.max_seg);
            lds = ((lds + ec + 7) & ~size_t(7)) + (size_t)a.max_img * (3 * sizeof(T) + 6 * sizeof(int)) +
                  (size_t)a.nproj * (2 * sizeof(T) + 2 * sizeof(int)) + (size_t)(p.tz * p.ty + 1) * sizeof(int) +
                  (size_t)a.max_seg * sizeof(uint32_t);
        } else {
            const size_t ring = (size_t)p.nb * p.ty * g.n[0] * sizeof(T);
            // fewer planes per projector phase when the accumulator ring would not fit
            while (p.nlp > 1 && lds + p.nlp * ring + ec > 65536) p.nlp /= 2;
            a.nlp = p.nlp;
            lds += p.nlp * ring + ec;
        }
    }
    QG_REQUIRE(lds <= 65536, "tile staging does not fit in 64 KiB of LDS");
    a.tile_major = (!a.has_nl && MODE != MODE_NL && env_int("QG_TILE_MAJOR", 1) != 0) ? 1 : 0;
    const StencilCoef<T> sc = make_coef<T>(g, coef);
    const int nchunk = (nbands + p.nb - 1) / p.nb;
    const int64_t nwg = (int64_t)a.ntile_y * a.nslab * nchunk;
    QG_REQUIRE(nwg < (1LL << 31), "too many workgroups");
    return FusedSetup<T>{p, a, sc, lds, njs, nwg};
}

template <class T, int MODE>
static void launch_fused(const qg_grid& g, int nbands, const T* psi, int64_t ld_in, T* out,
                         int64_t ld_out, double coef, int accumulate, const T* V,
                         qg_nonlocal* nl, const double* coefs, hipStream_t s) {
    const FusedSetup<T> fs = setup_fused<T, MODE>(g, nbands, ld_in, ld_out, coef, accumulate, nl, coefs);
    const Plan& p = fs.p;
    const FusedArgs& a = fs.a;
    const StencilCoef<T>& sc = fs.sc;
    const size_t lds = fs.lds;
    const int njs = fs.njs;
    dim3 grid((unsigned)fs.nwg);
    if constexpr (MODE == MODE_NL) {
        if (p.nb >= 4) {
            switch (p.nb) {
                case 4:
                    if (njs == 4) launch_fused_ppt<T, 4, MODE, 4>(p, grid, lds, s, a, sc, V, psi, out);
                    else launch_fused_ppt<T, 4, MODE, 8>(p, grid, lds, s, a, sc, V, psi, out);
                    break;
                case 8:
                    if (njs == 4) launch_fused_ppt<T, 8, MODE, 4>(p, grid, lds, s, a, sc, V, psi, out);
                    else launch_fused_ppt<T, 8, MODE, 8>(p, grid, lds, s, a, sc, V, psi, out);
                    break;
                default:
                    if (njs == 4) launch_fused_ppt<T, 16, MODE, 4>(p, grid, lds, s, a, sc, V, psi, out);
                    else launch_fused_ppt<T, 16, MODE, 8>(p, grid, lds, s, a, sc, V, psi, out);
                    break;
            }
            return;
        }
    }
    if (p.nb == 2) {
        if (njs == 4) launch_fused_ppt<T, 2, MODE, 4>(p, grid, lds, s, a, sc, V, psi, out);
        else launch_fused_ppt<T, 2, MODE, 8>(p, grid, lds, s, a, sc, V, psi, out);
    } else {
        if (njs == 4) launch_fused_ppt<T, 1, MODE, 4>(p, grid, lds, s, a, sc, V, psi, out);
        else launch_fused_ppt<T, 1, MODE, 8>(p, grid, lds, s, a, sc, V, psi, out);
    }
}

template <class T>
static void launch_project(qg_nonlocal* nl, int nbands, const T* psi, int64_t ld, double* coef,
                           hipStream_t s) {
    ProjArgs a{};
    a.natoms = nl->geo.natoms;
    a.nbands = nbands;
    a.ld = ld;
    a.dv = nl->gcopy.dvol();
    a.atom_img_off = nl->d_atom_img_off.p;
    a.atom_coef_off = nl->d_atom_coef_off.p;
    a.img_tbl_off = nl->d_img_tbl_off.p;
    a.gidx = nl->d_gidx.p;
    a.beta = (sizeof(T) == 4) ? (const void*)nl->d_beta32.p : (const void*)nl->d_beta64.p;
    a.coef = coef;
    a.ncoef = nl->geo.ncoef;
    const int nba = env_int("QG_PROJ_NBA", 8);
    if (nl->geo.njs == 4) {
        auto go = [&](auto nba_c) {
            constexpr int NBA = decltype(nba_c)::value;
            const int nchunk = (nbands + NBA - 1) / NBA;
            hipLaunchKernelGGL((project_kernel<T, NBA, 4>), dim3((unsigned)((int64_t)a.natoms * nchunk)),
                               dim3(kProjBlock), 0, s, a, psi);
        };
        if (nba >= 16 && nbands >= 16) go(std::integral_constant<int, 16>{});
        else if (nba >= 8) go(std::integral_constant<int, 8>{});
        else go(std::integral_constant<int, 4>{});
    } else
        hipLaunchKernelGGL((project_kernel<T, 4, 8>), dim3((unsigned)((int64_t)a.natoms * ((nbands + 3) / 4))),
                           dim3(kProjBlock), 0, s, a, psi);
    QG_HIP(hipGetLastError());
}

// Bands per chunk so that one chunk of psi stays resident in the Infinity Cache between
// the projection and the fused apply (QG_CHUNK_MB, default 128 MB).
static int chunk_bands(const qg_grid& g, int dtype, int nbands, bool split) {
    const int64_t bytes = g.ng() * (dtype == QG_F32 ? 4 : 8);
    // the split form keeps nothing in the Infinity Cache between its passes, so its chunks
    // only bound the coefficient scratch
    const int64_t budget = (int64_t)env_int("QG_CHUNK_MB", split ? 512 : 128) << 20;
    int64_t c = std::max<int64_t>(1, budget / std::max<int64_t>(1, bytes));
    // whole groups of the projection's 8 bands per workgroup (9-band chunks left a 1-band group)
    if (c > 8) c = c / 8 * 8;
    return (int)std::min<int64_t>(nbands, c);
}

// H psi for a band block, used by qg_hamiltonian_apply and the LOBPCG driver.
void hamiltonian_apply(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi,
                       int64_t ld_out, hipStream_t s) {
    // the row kernels (one projection pass, one stencil + V + projector pass) where they
    // serve the grid and projector set; the table kernels below otherwise
    if (rows_hamiltonian_apply(h, dtype, nbands, psi, ld, hpsi, ld_out, s, nullptr)) return;
    const qg_grid& g = h->grid;
    qg_nonlocal* nl = (h->nl && h->nl->geo.ncoef > 0) ? h->nl : nullptr;
    // split form (stencil + V pass, then a projector pass over many bands at a time) for FP32
    // with in-kernel projector values; one fused pass otherwise (QG_SPLIT_NL=0/1 forces)
    const int split_env = env_int("QG_SPLIT_NL", -1);
    const bool split = nl && (split_env >= 0 ? split_env != 0 : segments_form(nl, dtype, MODE_NL));
    const int cb = nl ? chunk_bands(g, dtype, nbands, split) : nbands;
    if (nl) nl->d_coef.alloc((size_t)std::min(cb, nbands) * nl->geo.ncoef);
    const int nchunks = (nbands + cb - 1) / cb;
    if (h->timing) {
        while ((int)h->tev.size() < 3 * nchunks) {
            hipEvent_t e;
            QG_HIP(hipEventCreate(&e));
            h->tev.push_back(e);
        }
        h->tchunks = nchunks;
    }
    for (int b0 = 0; b0 < nbands; b0 += cb) {
        const int ci = b0 / cb;
        if (h->timing) QG_HIP(hipEventRecord(h->tev[3 * ci], s));
        const int nb = std::min(cb, nbands - b0);
        const size_t es = (dtype == QG_F32) ? 4 : 8;
        const char* pin = (const char*)psi + (size_t)b0 * ld * es;
        char* pout = (char*)hpsi + (size_t)b0 * ld_out * es;
        if (nl) {
            if (dtype == QG_F64) launch_project<double>(nl, nb, (const double*)pin, ld, nl->d_coef.p, s);
            else launch_project<float>(nl, nb, (const float*)pin, ld, nl->d_coef.p, s);
        }
        if (h->timing) QG_HIP(hipEventRecord(h->tev[3 * ci + 1], s));
        if (nl && split) {
            // stencil + V pass, then the projector term as a read-modify-write pass over the
            // chunk (still resident in the Infinity Cache) with many bands per item
            if (dtype == QG_F64) {
                launch_fused<double, MODE_HAM>(g, nb, (const double*)pin, ld, (double*)pout, ld_out, -0.5, 0,
                                               h->V64.p, nullptr, nullptr, s);
                launch_fused<double, MODE_NL>(g, nb, (const double*)pin, ld, (double*)pout, ld_out, 0.0, 1,
                                              nullptr, nl, nl->d_coef.p, s);
            } else {
                launch_fused<float, MODE_HAM>(g, nb, (const float*)pin, ld, (float*)pout, ld_out, -0.5, 0,
                                              h->V32.p, nullptr, nullptr, s);
                launch_fused<float, MODE_NL>(g, nb, (const float*)pin, ld, (float*)pout, ld_out, 0.0, 1, nullptr,
                                             nl, nl->d_coef.p, s);
            }
        } else if (dtype == QG_F64) {
            launch_fused<double, MODE_HAM>(g, nb, (const double*)pin, ld, (double*)pout, ld_out, -0.5, 0,
                                           h->V64.p, nl, nl ? nl->d_coef.p : nullptr, s);
        } else {
            launch_fused<float, MODE_HAM>(g, nb, (const float*)pin, ld, (float*)pout, ld_out, -0.5, 0,
                                          h->V32.p, nl, nl ? nl->d_coef.p : nullptr, s);
        }
        if (h->timing) QG_HIP(hipEventRecord(h->tev[3 * ci + 2], s));
    }
    if (h->timing) h->timed_valid = true;
}

// ------------------------------------------------------------ batches of cells ----
// H psi for a batch of independent cells (different grids, atoms and band counts) in
// persistent launches: per round, one projection launch over every (cell, atom, band chunk)
// and one fused launch over every (cell, tile, slab, band) workgroup.  Rounds follow each
// cell's Infinity-Cache chunking of its band block.
struct BatchReq {
    qg_ham* h;
    int nbands;
    const void* psi;
    void* out;
};

template <class T, int NJS>
static void launch_fused_batch(int ppt, int nthr, size_t lds, int total, const BatchCell<T>* cells,
                               const int* prefix, int ncells, hipStream_t s) {
    const int grid = std::max(1, std::min(total, 2048));
    switch (ppt) {
#define QG_CASE(PP)                                                                                          \
    case PP:                                                                                                 \
        hipLaunchKernelGGL((fused_batch_kernel<T, 1, PP, MODE_HAM, NJS>), dim3(grid), dim3(nthr), lds, s, cells, \
                           prefix, ncells);                                                                  \
        break;
        QG_CASE(2) QG_CASE(4) QG_CASE(5) QG_CASE(8)
#undef QG_CASE
        default: throw Error{QG_ERR_ARG, "unsupported points-per-thread"};
    }
    QG_HIP(hipGetLastError());
}

// Batched projector passes (split form): NB bands per workgroup and PPT points per thread are
// shared by the cells of one launch; the host groups cells by them.
template <class T, int NJS>
static void launch_nl_batch(int nb, int ppt, int nthr, size_t lds, int total, const BatchCell<T>* cells,
                            const int* prefix, int ncells, hipStream_t s) {
    const int grid = std::max(1, std::min(total, 2048));
#define QG_NLB(NBV, PP)                                                                                     \
    hipLaunchKernelGGL((fused_batch_kernel<T, NBV, PP, MODE_NL, NJS>), dim3(grid), dim3(nthr), lds, s, cells, \
                       prefix, ncells)
    if (ppt == 2) {
        switch (nb) {
            case 8: QG_NLB(8, 2); break;
            case 4: QG_NLB(4, 2); break;
            case 2: QG_NLB(2, 2); break;
            default: QG_NLB(1, 2); break;
        }
    } else {
        switch (nb) {
            case 8: QG_NLB(8, 4); break;
            case 4: QG_NLB(4, 4); break;
            case 2: QG_NLB(2, 4); break;
            default: QG_NLB(1, 4); break;
        }
    }
#undef QG_NLB
    QG_HIP(hipGetLastError());
}

template <class T>
static void batch_apply_t(const std::vector<BatchReq>& reqs, hipStream_t s, DevBuf<unsigned char>& scratch,
                          DevBuf<double>& coefs) {
    const int dtype = sizeof(T) == 4 ? QG_F32 : QG_F64;
    const int n = (int)reqs.size();
    std::vector<int> cb(n);
    int rounds = 0;
    int njs = 0;
    // split form (stencil + V launch, then projector-pass launches) when every cell with
    // projectors would use it on its own
    const int split_env = env_int("QG_SPLIT_NL", -1);
    bool split = false, all_split = true;
    for (int c = 0; c < n; ++c) {
        qg_nonlocal* nl = (reqs[c].h->nl && reqs[c].h->nl->geo.ncoef > 0) ? reqs[c].h->nl : nullptr;
        if (!nl) continue;
        const bool sc = split_env >= 0 ? split_env != 0 : segments_form(nl, dtype, MODE_NL);
        split = split || sc;
        all_split = all_split && sc;
    }
    split = split && all_split;
    // every request gets its own coefficient region (cells may share one nonlocal handle,
    // e.g. two spin channels on one integrator)
    std::vector<size_t> coff(n, 0);
    size_t ctotal = 0;
    for (int c = 0; c < n; ++c) {
        const BatchReq& r = reqs[c];
        qg_nonlocal* nl = (r.h->nl && r.h->nl->geo.ncoef > 0) ? r.h->nl : nullptr;
        cb[c] = nl ? chunk_bands(r.h->grid, dtype, r.nbands, split) : std::max(1, r.nbands);
        if (nl) {
            coff[c] = ctotal;
            ctotal += (size_t)std::min(cb[c], r.nbands) * nl->geo.ncoef;
            QG_REQUIRE(njs == 0 || njs == nl->geo.njs, "cells of one batch must share the projector slot count");
            njs = nl->geo.njs;
        }
        rounds = std::max(rounds, (r.nbands + cb[c] - 1) / cb[c]);
    }
    coefs.alloc(std::max<size_t>(1, ctotal));
    if (njs == 0) njs = 4;
    const int NBA = njs == 4 ? 8 : 4;
    // descriptors of every round, packed into one upload
    struct RoundDesc {
        size_t pc, pp;
        int nproj, proj_total;
    };
    std::vector<RoundDesc> rd(rounds);
    // stencil cells grouped by points per thread (the autotuner may pick it per grid shape)
    struct FusedGroupDesc {
        size_t fc, fp;
        int ppt, ncells, total, nthr;
        size_t lds;
    };
    std::vector<std::vector<FusedGroupDesc>> fgd(rounds);
    struct NlGroupDesc {
        size_t fc, fp;
        int nb, ppt, ncells, total, nthr;
        size_t lds;
    };
    std::vector<std::vector<NlGroupDesc>> nlg(rounds);
    std::vector<unsigned char> host;
    auto push = [&](const void* src, size_t bytes) {
        const size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + bytes);
        std::memcpy(host.data() + off, src, bytes);
        return off;
    };
    for (int r = 0; r < rounds; ++r) {
        std::vector<ProjCell<T>> pcs;
        std::vector<int> ppre{0};
        struct FusedGroup {
            std::vector<BatchCell<T>> cells;
            std::vector<int> pre{0};
            int nthr = 0;
            size_t lds = 0;
        };
        std::map<int, FusedGroup> fgroups;
        struct NlGroup {
            std::vector<BatchCell<T>> cells;
            std::vector<int> pre{0};
            int nthr = 0;
            size_t lds = 0;
        };
        std::map<std::pair<int, int>, NlGroup> groups;
        for (int c = 0; c < n; ++c) {
            const BatchReq& q = reqs[c];
            const int b0 = r * cb[c];
            if (b0 >= q.nbands) continue;
            const int nb = std::min(cb[c], q.nbands - b0);
            const qg_grid& g = q.h->grid;
            const int64_t ng = g.ng();
            const T* pin = (const T*)q.psi + (int64_t)b0 * ng;
            T* pout = (T*)q.out + (int64_t)b0 * ng;
            qg_nonlocal* nl = (q.h->nl && q.h->nl->geo.ncoef > 0) ? q.h->nl : nullptr;
            if (nl) {
                ProjArgs a{};
                a.natoms = nl->geo.natoms;
                a.nbands = nb;
                a.ld = ng;
                a.dv = nl->gcopy.dvol();
                a.atom_img_off = nl->d_atom_img_off.p;
                a.atom_coef_off = nl->d_atom_coef_off.p;
                a.img_tbl_off = nl->d_img_tbl_off.p;
                a.gidx = nl->d_gidx.p;
                a.beta = (sizeof(T) == 4) ? (const void*)nl->d_beta32.p : (const void*)nl->d_beta64.p;
                a.coef = coefs.p + coff[c];
                a.ncoef = nl->geo.ncoef;
                pcs.push_back(ProjCell<T>{a, pin});
                ppre.push_back(ppre.back() + a.natoms * ((nb + NBA - 1) / NBA));
            }
            qg_nonlocal* fnl = split ? nullptr : nl;
            double* ccoef = nl ? coefs.p + coff[c] : nullptr;
            const FusedSetup<T> fs = setup_fused<T, MODE_HAM>(g, nb, ng, ng, -0.5, 0, fnl, fnl ? ccoef : nullptr);
            QG_REQUIRE(fs.p.nb == 1, "batched cells use one band per workgroup");
            if (split && nl) {
                const FusedSetup<T> fn = setup_fused<T, MODE_NL>(g, nb, ng, ng, 0.0, 1, nl, ccoef);
                QG_REQUIRE(fn.p.nb == 1 || fn.p.nb == 2 || fn.p.nb == 4 || fn.p.nb == 8,
                           "batched projector pass: unsupported bands per workgroup");
                QG_REQUIRE(fn.p.ppt == 2 || fn.p.ppt == 4, "batched projector pass: unsupported points per thread");
                NlGroup& gr = groups[{fn.p.nb, fn.p.ppt}];
                gr.cells.push_back(BatchCell<T>{fn.a, fn.sc, nullptr, pin, pout});
                QG_REQUIRE((int64_t)gr.pre.back() + fn.nwg < (1LL << 31), "batch too large");
                gr.pre.push_back(gr.pre.back() + (int)fn.nwg);
                gr.nthr = std::max(gr.nthr, fn.p.nthr);
                gr.lds = std::max(gr.lds, fn.lds);
            }
            FusedGroup& fg = fgroups[fs.p.ppt];
            fg.nthr = std::max(fg.nthr, fs.p.nthr);
            fg.lds = std::max(fg.lds, fs.lds);
            const T* V = (sizeof(T) == 4) ? (const T*)q.h->V32.p : (const T*)q.h->V64.p;
            fg.cells.push_back(BatchCell<T>{fs.a, fs.sc, V, pin, pout});
            QG_REQUIRE((int64_t)fg.pre.back() + fs.nwg < (1LL << 31), "batch too large");
            fg.pre.push_back(fg.pre.back() + (int)fs.nwg);
        }
        RoundDesc& d = rd[r];
        d.nproj = (int)pcs.size();
        d.proj_total = ppre.back();
        d.pc = pcs.empty() ? 0 : push(pcs.data(), pcs.size() * sizeof(ProjCell<T>));
        d.pp = push(ppre.data(), ppre.size() * sizeof(int));
        for (auto& kv : fgroups) {
            FusedGroup& fg = kv.second;
            FusedGroupDesc gd{};
            gd.ppt = kv.first;
            gd.ncells = (int)fg.cells.size();
            gd.total = fg.pre.back();
            gd.nthr = fg.nthr;
            gd.lds = fg.lds;
            gd.fc = push(fg.cells.data(), fg.cells.size() * sizeof(BatchCell<T>));
            gd.fp = push(fg.pre.data(), fg.pre.size() * sizeof(int));
            fgd[r].push_back(gd);
        }
        for (auto& kv : groups) {
            NlGroup& gr = kv.second;
            NlGroupDesc gd{};
            gd.nb = kv.first.first;
            gd.ppt = kv.first.second;
            gd.ncells = (int)gr.cells.size();
            gd.total = gr.pre.back();
            gd.nthr = gr.nthr;
            gd.lds = gr.lds;
            gd.fc = push(gr.cells.data(), gr.cells.size() * sizeof(BatchCell<T>));
            gd.fp = push(gr.pre.data(), gr.pre.size() * sizeof(int));
            nlg[r].push_back(gd);
        }
    }
    scratch.alloc(host.size());
    QG_HIP(hipMemcpyAsync(scratch.p, host.data(), host.size(), hipMemcpyHostToDevice, s));
    for (int r = 0; r < rounds; ++r) {
        const RoundDesc& d = rd[r];
        const unsigned char* b = scratch.p;
        if (d.nproj > 0 && d.proj_total > 0) {
            const int grid = std::max(1, std::min(d.proj_total, 4096));
            if (njs == 4)
                hipLaunchKernelGGL((project_batch_kernel<T, 8, 4>), dim3(grid), dim3(kProjBlock), 0, s,
                                   (const ProjCell<T>*)(b + d.pc), (const int*)(b + d.pp), d.nproj);
            else
                hipLaunchKernelGGL((project_batch_kernel<T, 4, 8>), dim3(grid), dim3(kProjBlock), 0, s,
                                   (const ProjCell<T>*)(b + d.pc), (const int*)(b + d.pp), d.nproj);
            QG_HIP(hipGetLastError());
        }
        for (const FusedGroupDesc& gd : fgd[r]) {
            if (gd.total <= 0) continue;
            if (njs == 4)
                launch_fused_batch<T, 4>(gd.ppt, gd.nthr, gd.lds, gd.total, (const BatchCell<T>*)(b + gd.fc),
                                         (const int*)(b + gd.fp), gd.ncells, s);
            else
                launch_fused_batch<T, 8>(gd.ppt, gd.nthr, gd.lds, gd.total, (const BatchCell<T>*)(b + gd.fc),
                                         (const int*)(b + gd.fp), gd.ncells, s);
        }
        for (const NlGroupDesc& gd : nlg[r]) {
            if (gd.total <= 0) continue;
            if (njs == 4)
                launch_nl_batch<T, 4>(gd.nb, gd.ppt, gd.nthr, gd.lds, gd.total, (const BatchCell<T>*)(b + gd.fc),
                                      (const int*)(b + gd.fp), gd.ncells, s);
            else
                launch_nl_batch<T, 8>(gd.nb, gd.ppt, gd.nthr, gd.lds, gd.total, (const BatchCell<T>*)(b + gd.fc),
                                      (const int*)(b + gd.fp), gd.ncells, s);
        }
    }
    // the host descriptors must outlive the upload
    QG_HIP(hipStreamSynchronize(s));
}

void batch_hamiltonian_apply(const std::vector<BatchReq>& reqs, int dtype, hipStream_t s,
                             DevBuf<unsigned char>& scratch, DevBuf<double>& coefs) {
    if (reqs.empty()) return;
    if (dtype == QG_F64) batch_apply_t<double>(reqs, s, scratch, coefs);
    else batch_apply_t<float>(reqs, s, scratch, coefs);
}

// ----------------------------------------------------------------- autotuner ----
// Times the fused H psi of this Hamiltonian for candidate tile shapes (rows per tile and
// points per thread; slab depth by the planner's rule) on zero-filled blocks of nbands bands
// and records the fastest for the grid shape and dtype.
static void autotune(qg_ham* h, int dtype, int nbands, int fixed_ppt, TunedShape& best, double& best_ms,
                     hipStream_t s) {
    const qg_grid& g = h->grid;
    const int nx = g.n[0], ny = g.n[1];
    const size_t es = dtype == QG_F32 ? 4 : 8;
    const int64_t ng = g.ng();
    DevBuf<unsigned char> in, out;
    in.alloc((size_t)nbands * ng * es);
    out.alloc((size_t)nbands * ng * es);
    QG_HIP(hipMemsetAsync(in.p, 0, (size_t)nbands * ng * es, s));
    hipEvent_t e0, e1;
    QG_HIP(hipEventCreate(&e0));
    QG_HIP(hipEventCreate(&e1));
    std::vector<int> ppts;
    if (fixed_ppt) ppts = {fixed_ppt};
    else if (dtype == QG_F32) ppts = {2, 4};
    else ppts = {2, 4};
    best_ms = 1e30;
    for (int ppt : ppts)
        for (int ty : {4, 6, 8, 12, 16, 20, 24, 32}) {
            if (ty > ny) continue;
            const int nthr = round64((ty * nx + ppt - 1) / ppt);
            if (nthr > kMaxThreads || std::max(nthr, round64((8 * nx + kHaloPT - 1) / kHaloPT)) > kMaxThreads) continue;
            TunedShape trial{ty, ppt};
            t_trial = &trial;
            try {
                hamiltonian_apply(h, dtype, nbands, in.p, ng, out.p, ng, s);   // warm-up (+ tables)
                QG_HIP(hipEventRecord(e0, s));
                for (int r = 0; r < 3; ++r) hamiltonian_apply(h, dtype, nbands, in.p, ng, out.p, ng, s);
                QG_HIP(hipEventRecord(e1, s));
                QG_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                QG_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms / 3 < best_ms) {
                    best_ms = ms / 3;
                    best = trial;
                }
            } catch (const Error& e) {
                if (e.code != QG_ERR_ARG) {
                    t_trial = nullptr;
                    (void)hipEventDestroy(e0);
                    (void)hipEventDestroy(e1);
                    throw;
                }
            }
            t_trial = nullptr;
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    QG_REQUIRE(best_ms < 1e30, "no feasible tile shape");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[std::make_tuple(g.n[0], g.n[1], g.n[2], dtype)] = best;
}

// (T + V) psi: the local part of H (no projectors), used by the learned preconditioner's
// training targets.
void local_hamiltonian_apply(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* out,
                             int64_t ld_out, hipStream_t s) {
    const qg_grid& g = h->grid;
    if (dtype == QG_F64)
        launch_fused<double, MODE_HAM>(g, nbands, (const double*)psi, ld, (double*)out, ld_out, -0.5, 0, h->V64.p,
                                       nullptr, nullptr, s);
    else
        launch_fused<float, MODE_HAM>(g, nbands, (const float*)psi, ld, (float*)out, ld_out, -0.5, 0, h->V32.p,
                                      nullptr, nullptr, s);
}

}  // namespace qg

using namespace qg;

static void check_block(const qg_grid* g, int dtype, int nbands, const void* psi, int64_t ld,
                        const void* out, int64_t ld_out) {
    QG_REQUIRE(g, "null grid");
    QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "dtype must be QG_F64 or QG_F32");
    QG_REQUIRE(nbands >= 0, "negative band count");
    QG_REQUIRE(nbands == 0 || (psi && out), "null psi/out");
    QG_REQUIRE(ld >= g->ng() && ld_out >= g->ng(), "leading dimension smaller than the grid");
    QG_REQUIRE(psi != out, "in-place application is not supported (psi must differ from out)");
}

// (Re)build the projector geometry of nl for natoms nuclei with the channels saved at
// creation, in place: the handle keeps its address, so Hamiltonians built on it see the new
// positions at their next application (the reference's UpdatePosition on a live integrator,
// PseudoPotOperator_mpi_v2.h:231-248).  Cached per-plan staging is dropped.
static void nl_build(qg_nonlocal* nl, int natoms, const double* xyz, const int* species) {
    for (int a = 0; a < natoms; ++a)
        QG_REQUIRE(species[a] >= 0 && species[a] < nl->nspecies, "species index out of range");
    NlGeometry geo;
    build_nl_geometry(nl->gcopy, natoms, xyz, species, nl->nspecies, nl->nproj.data(), nl->specs.data(), geo);
    nl->geo = std::move(geo);
    ++nl->geo_version;
    const auto& gm = nl->geo;
    if (!gm.projs.empty()) nl->d_projs.upload(gm.projs.data(), gm.projs.size());
    if (!gm.tables.empty()) nl->d_tables.upload(gm.tables.data(), gm.tables.size());
    if (!gm.images.empty()) nl->d_images.upload(gm.images.data(), gm.images.size());
    if (!gm.intervals.empty()) nl->d_intervals.upload(gm.intervals.data(), gm.intervals.size());
    if (!gm.gidx.empty()) {
        nl->d_gidx.upload(gm.gidx.data(), gm.gidx.size());
        nl->d_beta64.upload(gm.beta.data(), gm.beta.size());
        std::vector<float> b32(gm.beta.begin(), gm.beta.end());
        nl->d_beta32.upload(b32.data(), b32.size());
    }
    nl->d_img_tbl_off.upload(gm.img_tbl_off.data(), gm.img_tbl_off.size());
    if (!gm.coef_energy.empty()) nl->d_coef_energy.upload(gm.coef_energy.data(), gm.coef_energy.size());
    nl->d_atom_img_off.upload(gm.atom_img_off.data(), gm.atom_img_off.size());
    nl->d_atom_coef_off.upload(gm.atom_coef_off.data(), gm.atom_coef_off.size());
    for (auto& t : nl->tiles) t.ty = t.tz = -1;
    nl->rows.clear();
    rows_prepare(nl);
    // Gaussian projectors can be evaluated inside the projector pass instead of read from
    // per-point tables (QG_NL_ANALYTIC=0 keeps the tables everywhere)
    bool gauss = !gm.projs.empty();
    for (const DevProj& d : gm.projs) gauss = gauss && d.kind == QG_RADIAL_GAUSSIAN && d.l <= 1;
    nl->analytic_level = env_int("QG_NL_ANALYTIC", 1);
    nl->analytic = gauss && nl->gcopy.n[0] <= 1024 && nl->analytic_level != 0;
}

extern "C" {

int qg_laplacian8_apply(const qg_grid* g, int dtype, int nbands, const void* psi, int64_t ld_psi,
                        void* out, int64_t ld_out, double coef, int accumulate, void* stream) {
    return guarded([&] {
        check_block(g, dtype, nbands, psi, ld_psi, out, ld_out);
        if (nbands == 0) return;
        hipStream_t s = as_stream(stream);
        if (rows_laplacian_apply(*g, dtype, nbands, psi, ld_psi, out, ld_out, coef, accumulate, s)) return;
        if (dtype == QG_F64)
            launch_fused<double, MODE_LAP>(*g, nbands, (const double*)psi, ld_psi, (double*)out, ld_out, coef,
                                           accumulate, nullptr, nullptr, nullptr, s);
        else
            launch_fused<float, MODE_LAP>(*g, nbands, (const float*)psi, ld_psi, (float*)out, ld_out, coef,
                                          accumulate, nullptr, nullptr, nullptr, s);
    });
}

int qg_nonlocal_create(const qg_grid* g, int natoms, const double* xyz, const int* species,
                       int nspecies, const int* nproj, const qg_proj_spec* specs, qg_nonlocal** out) {
    return guarded([&] {
        QG_REQUIRE(g && out, "null argument");
        QG_REQUIRE(natoms >= 0 && nspecies >= 0, "negative counts");
        QG_REQUIRE(natoms == 0 || (xyz && species && nproj && (specs || nspecies == 0)), "null arrays");
        qg_nonlocal* nl = new qg_nonlocal;
        try {
            nl->gcopy = *g;
            nl->grid = &nl->gcopy;
            nl->nspecies = nspecies;
            int total = 0;
            for (int sp = 0; sp < nspecies; ++sp) {
                QG_REQUIRE(nproj[sp] >= 0, "negative projector count");
                nl->nproj.push_back(nproj[sp]);
                total += nproj[sp];
            }
            nl->specs.assign(specs, specs + total);
            nl->spec_tables.resize(total);
            for (int k = 0; k < total; ++k)
                if (specs[k].radial_kind == QG_RADIAL_TABLE && specs[k].table && specs[k].n_table > 0) {
                    nl->spec_tables[k].assign(specs[k].table, specs[k].table + specs[k].n_table);
                    nl->specs[k].table = nl->spec_tables[k].data();
                }
            nl_build(nl, natoms, xyz, species);
        } catch (...) {
            delete nl;
            throw;
        }
        *out = nl;
    });
}

int qg_nonlocal_update(qg_nonlocal* nl, int natoms, const double* xyz, const int* species) {
    return guarded([&] {
        QG_REQUIRE(nl, "null nonlocal handle");
        QG_REQUIRE(natoms >= 0, "negative atom count");
        QG_REQUIRE(natoms == 0 || (xyz && species), "null arrays");
        // device work that may still read the old geometry finishes first
        QG_HIP(hipDeviceSynchronize());
        nl_build(nl, natoms, xyz, species);
    });
}

int qg_nonlocal_destroy(qg_nonlocal* nl) {
    delete nl;
    return QG_OK;
}

int qg_nonlocal_num_coef(const qg_nonlocal* nl) { return nl ? nl->geo.ncoef : -1; }

int qg_nonlocal_projector_form(const qg_nonlocal* nl) { return nl ? (nl->analytic ? 1 : 0) : -1; }

int qg_nonlocal_project(qg_nonlocal* nl, int dtype, int nbands, const void* psi, int64_t ld,
                        double* coef, void* stream) {
    return guarded([&] {
        QG_REQUIRE(nl, "null nonlocal handle");
        QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "bad dtype");
        QG_REQUIRE(ld >= nl->gcopy.ng(), "leading dimension smaller than the grid");
        if (nbands == 0 || nl->geo.ncoef == 0) return;
        QG_REQUIRE(psi && coef, "null pointer");
        hipStream_t s = as_stream(stream);
        if (rows_nonlocal_project(nl, dtype, nbands, psi, ld, coef, nl->rows_part, s)) return;
        if (dtype == QG_F64)
            launch_project<double>(nl, nbands, (const double*)psi, ld, coef, s);
        else
            launch_project<float>(nl, nbands, (const float*)psi, ld, coef, s);
    });
}

int qg_nonlocal_apply(qg_nonlocal* nl, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi,
                      int64_t ld_out, int accumulate, void* stream) {
    return guarded([&] {
        QG_REQUIRE(nl, "null nonlocal handle");
        check_block(&nl->gcopy, dtype, nbands, psi, ld, hpsi, ld_out);
        if (nbands == 0) return;
        hipStream_t s = as_stream(stream);
        if (nl->geo.ncoef == 0) {
            // no projectors: out = accumulate ? out : 0, through the same kernel
            if (dtype == QG_F64)
                launch_fused<double, MODE_NL>(nl->gcopy, nbands, (const double*)psi, ld, (double*)hpsi, ld_out,
                                              0.0, accumulate, nullptr, nullptr, nullptr, s);
            else
                launch_fused<float, MODE_NL>(nl->gcopy, nbands, (const float*)psi, ld, (float*)hpsi, ld_out, 0.0,
                                             accumulate, nullptr, nullptr, nullptr, s);
            return;
        }
        nl->d_coef.alloc((size_t)nbands * nl->geo.ncoef);
        if (dtype == QG_F64) {
            launch_project<double>(nl, nbands, (const double*)psi, ld, nl->d_coef.p, s);
            launch_fused<double, MODE_NL>(nl->gcopy, nbands, (const double*)psi, ld, (double*)hpsi, ld_out, 0.0,
                                          accumulate, nullptr, nl, nl->d_coef.p, s);
        } else {
            launch_project<float>(nl, nbands, (const float*)psi, ld, nl->d_coef.p, s);
            launch_fused<float, MODE_NL>(nl->gcopy, nbands, (const float*)psi, ld, (float*)hpsi, ld_out, 0.0,
                                         accumulate, nullptr, nl, nl->d_coef.p, s);
        }
    });
}

int qg_hamiltonian_create(const qg_grid* g, const double* V, qg_nonlocal* nl, qg_ham** out) {
    return guarded([&] {
        QG_REQUIRE(g && V && out, "null argument");
        if (nl)
            QG_REQUIRE(nl->gcopy.n[0] == g->n[0] && nl->gcopy.n[1] == g->n[1] && nl->gcopy.n[2] == g->n[2],
                       "nonlocal handle was built for another grid");
        // device address of the allocation-free entry's scratch on this device (set-up, once per device)
        (void)scratch_address();
        qg_ham* h = new qg_ham;
        h->grid = *g;
        h->nl = nl;
        try {
            const int64_t ng = g->ng();
            h->V64.upload(V, ng);
            std::vector<float> v32(V, V + ng);
            h->V32.upload(v32.data(), ng);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int qg_hamiltonian_destroy(qg_ham* h) {
    delete h;
    return QG_OK;
}

int qg_hamiltonian_set_potential(qg_ham* h, const double* V) {
    return guarded([&] {
        QG_REQUIRE(h && V, "null argument");
        const int64_t ng = h->grid.ng();
        h->V64.upload(V, ng);
        std::vector<float> v32(V, V + ng);
        h->V32.upload(v32.data(), ng);
    });
}

int qg_hamiltonian_apply(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi,
                         int64_t ld_out, void* stream) {
    return guarded([&] {
        QG_REQUIRE(h, "null hamiltonian");
        check_block(&h->grid, dtype, nbands, psi, ld, hpsi, ld_out);
        if (nbands == 0) return;
        hamiltonian_apply(h, dtype, nbands, psi, ld, hpsi, ld_out, as_stream(stream));
    });
}

void qg_default_hpsi_plan(qg_hpsi_plan* plan) {
    if (plan) *plan = qg_hpsi_plan{QG_KERNELS_AUTO, 0, 0, 0, 0, 0};
}

int qg_hamiltonian_set_plan(qg_ham* h, const qg_hpsi_plan* plan) {
    return guarded([&] {
        QG_REQUIRE(h, "null hamiltonian");
        qg_hpsi_plan p{};
        if (plan) p = *plan;
        QG_REQUIRE(p.kernels == QG_KERNELS_AUTO || p.kernels == QG_KERNELS_ROWS || p.kernels == QG_KERNELS_TABLES,
                   "unknown kernel selection");
        QG_REQUIRE(p.rows_ty >= 0 && p.rows_ty <= 32 && p.rows_tz >= 0 && p.rows_tzp >= 0, "bad tile sizes");
        QG_REQUIRE(p.rows_bands == 0 || p.rows_bands == 1 || p.rows_bands == 2 || p.rows_bands == 4,
                   "bands per workgroup must be 1, 2 or 4");
        QG_REQUIRE(p.rows_nbp == 0 || p.rows_nbp == 4 || p.rows_nbp == 8, "projection bands must be 4 or 8");
        h->plan = p;
    });
}

int qg_hamiltonian_apply_noalloc(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi,
                                 int64_t ld_out, void* stream) {
    return guarded([&] {
        QG_REQUIRE(h, "null hamiltonian");
        check_block(&h->grid, dtype, nbands, psi, ld, hpsi, ld_out);
        if (nbands == 0) return;
        rows_apply_noalloc(h, dtype, nbands, psi, ld, hpsi, ld_out, as_stream(stream));
    });
}

int qg_batch_hamiltonian_apply(int ncells, qg_ham* const* hams, int dtype, const int* nbands,
                               const void* const* psi, void* const* hpsi, void* stream) {
    return guarded([&] {
        QG_REQUIRE(ncells >= 0 && (ncells == 0 || (hams && nbands && psi && hpsi)), "bad arguments");
        QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "bad dtype");
        std::vector<BatchReq> reqs;
        for (int c = 0; c < ncells; ++c) {
            QG_REQUIRE(hams[c] && nbands[c] >= 0, "bad cell");
            if (nbands[c] == 0) continue;
            QG_REQUIRE(psi[c] && hpsi[c] && psi[c] != hpsi[c], "bad cell block");
            reqs.push_back(BatchReq{hams[c], nbands[c], psi[c], hpsi[c]});
        }
        // descriptor upload area of this entry point (kept for the life of the process)
        static DevBuf<unsigned char>* scratch = new DevBuf<unsigned char>();
        static DevBuf<double>* coefs = new DevBuf<double>();
        static std::mutex mu;
        std::lock_guard<std::mutex> lk(mu);
        batch_hamiltonian_apply(reqs, dtype, as_stream(stream), *scratch, *coefs);
    });
}

int qg_autotune_hamiltonian(qg_ham* h, int dtype, int nbands, int fixed_ppt, int* ty, int* ppt,
                            double* best_ms) {
    return guarded([&] {
        QG_REQUIRE(h && nbands >= 1, "bad arguments");
        QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "bad dtype");
        QG_REQUIRE(fixed_ppt == 0 || fixed_ppt == 2 || fixed_ppt == 4 || fixed_ppt == 5 || fixed_ppt == 8,
                   "fixed_ppt must be 0, 2, 4, 5 or 8");
        TunedShape best{};
        double ms = 0.0;
        autotune(h, dtype, nbands, fixed_ppt, best, ms, nullptr);
        if (ty) *ty = best.ty;
        if (ppt) *ppt = best.ppt;
        if (best_ms) *best_ms = ms;
    });
}

int qg_hamiltonian_set_timing(qg_ham* h, int enable) {
    return guarded([&] {
        QG_REQUIRE(h, "null hamiltonian");
        h->timing = enable != 0;
        h->timed_valid = false;
    });
}

int qg_hamiltonian_last_times(qg_ham* h, double* t_project_ms, double* t_fused_ms) {
    return guarded([&] {
        QG_REQUIRE(h && t_project_ms && t_fused_ms, "null argument");
        QG_REQUIRE(h->timing && h->timed_valid, "no timed apply recorded (enable timing first)");
        double ta = 0.0, tb = 0.0;
        for (int c = 0; c < h->tchunks; ++c) {
            QG_HIP(hipEventSynchronize(h->tev[3 * c + 2]));
            float a = 0.f, b = 0.f;
            QG_HIP(hipEventElapsedTime(&a, h->tev[3 * c], h->tev[3 * c + 1]));
            QG_HIP(hipEventElapsedTime(&b, h->tev[3 * c + 1], h->tev[3 * c + 2]));
            ta += a;
            tb += b;
        }
        *t_project_ms = ta;
        *t_fused_ms = tb;
    });
}

int qg_hamiltonian_last_chunks(qg_ham* h, int* nchunks) {
    return guarded([&] {
        QG_REQUIRE(h && nchunks, "null argument");
        QG_REQUIRE(h->timing && h->timed_valid, "no timed apply recorded (enable timing first)");
        *nchunks = h->tchunks;
    });
}

}  // extern "C"
