// qg_cnn.hip — learned preconditioner for the LOBPCG residual block.
//
// M r = K r + K N^T N K r, with K = (T + s)^-1 the exact kinetic preconditioner (qg_precond.hip)
// and N a 3-layer 3x3x3 convolution network on the periodic grid, 8 hidden channels, whose
// input is the residual u = K r together with the local-potential patch
//   x   = [u, v u]                         v = (V - mean V) / std V
//   h1  = conv(x, W1)  * g1,               g1_c = 1 + tanh(A1_c v + B1_c)
//   h2  = conv(h1, W2) * g2,               g2_c = 1 + tanh(A2_c v + B2_c)
//   N u = conv(h2, W3)
// N is linear in u (the potential enters through the gates), so M is a linear operator, and
// K N^T N K is positive semi-definite: M is symmetric positive definite for any weights, which
// keeps LOBPCG's Rayleigh-Ritz step well posed whatever the network has learned.  N^T is the
// exact adjoint (transposed convolutions with flipped taps, same gates).
//
// Training (on the device, by this file's kernels): pairs (r, z*) with z* = (T + V - sigma)^-1 r
// solved by preconditioned CG, r taken from the solver's own residual blocks; loss
// sum |K N^T N u - (z* - u)|^2 / sum |z*|^2 on rms-normalised samples; gradient by the chain
// rule dL = <w, dN g> + <N g, dN u>, g = 2 K e / Z, w = N u (two forward passes and two
// backward passes of N); Adam.  Weights are kept in FP32 for training; the preconditioner
// applies the BF16-rounded weights with FP32 accumulation.
//
// Convolution kernels: direct 3x3x3, one workgroup per 16 x 8 x 4 output tile and band, the
// input tile (+1 halo, periodic wrap) and the weights staged in LDS, each lane accumulating its
// 4 z-points x COUT channels in registers.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <vector>
#include "qg_internal.hpp"

namespace qg {

void precond_apply(qg_ham* h, int kind, double shift, int dtype, int nbands, void* r, int64_t ld, hipStream_t s);
void local_hamiltonian_apply(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* out,
                             int64_t ld_out, hipStream_t s);
void hamiltonian_apply(qg_ham* h, int dtype, int nbands, const void* psi, int64_t ld, void* hpsi,
                       int64_t ld_out, hipStream_t s);

constexpr int kC = 8;
constexpr int kTaps = 27;
constexpr int kOffW1 = 0;
constexpr int kOffW2 = kOffW1 + kC * 2 * kTaps;
constexpr int kOffW3 = kOffW2 + kC * kC * kTaps;
constexpr int kOffA1 = kOffW3 + kC * kTaps;
constexpr int kOffB1 = kOffA1 + kC;
constexpr int kOffA2 = kOffB1 + kC;
constexpr int kOffB2 = kOffA2 + kC;
constexpr int kNumParams = kOffB2 + kC;   // 2408

constexpr int kTX = 16, kTY = 8, kTZ = 4;
constexpr int kConvThreads = kTX * kTY;
constexpr int kLX = kTX + 2, kLY = kTY + 2, kLZ = kTZ + 2;
constexpr int kLCH = kLX * kLY * kLZ;

enum { IN_PLAIN = 0, IN_UV = 1 };
enum { OUT_PLAIN = 0, OUT_GATE = 1, OUT_UV = 2 };

struct ConvArgs {
    int nx, ny, nz;
    int64_t ng;
    int64_t in_bs, out_bs, pre_bs;   // band strides (elements)
    const float* v;                  // normalised potential
    const float* ga;                 // gate parameters A_c, B_c (COUT each)
    const float* gb;
};

// periodic index for any offset: a tile may overshoot the grid by up to a whole tile (kTX / kTY
// points) on grids smaller than the tile, so a single wrap is not enough
__device__ __forceinline__ int wrapi(int i, int n) { return ((i % n) + n) % n; }

// Stage the input tile (CIN channels, 1-point periodic halo) into LDS.
template <int CIN, int IN_MODE>
__device__ __forceinline__ void stage_tile(const ConvArgs& a, const float* inb, float* tile, int tx0, int ty0, int z0) {
    for (int i = threadIdx.x; i < kLCH; i += kConvThreads) {
        const int lx = i % kLX, ly = (i / kLX) % kLY, lz = i / (kLX * kLY);
        const int gx = wrapi(tx0 + lx - 1, a.nx), gy = wrapi(ty0 + ly - 1, a.ny), gz = wrapi(z0 + lz - 1, a.nz);
        const int64_t gi = gx + (int64_t)a.nx * (gy + (int64_t)a.ny * gz);
        if constexpr (IN_MODE == IN_UV) {
            const float u = inb[gi];
            tile[i] = u;
            tile[kLCH + i] = a.v[gi] * u;
        } else {
#pragma unroll
            for (int c = 0; c < CIN; ++c) tile[c * kLCH + i] = inb[c * a.ng + gi];
        }
    }
}

// out[b][co] = post( sum_{ci,tap} w[co][ci][tap] x[b][ci](p + d_tap) ), tap = 9 dz + 3 dy + dx
// over (dz, dy, dx) in {0,1,2}^3 meaning offsets -1..1.
template <int CIN, int COUT, int IN_MODE, int OUT_MODE, class WT>
__global__ __launch_bounds__(kConvThreads) void conv3_kernel(ConvArgs a, const WT* __restrict__ w,
                                                             const float* __restrict__ in, float* __restrict__ out,
                                                             float* __restrict__ pre) {
    __shared__ float tile[CIN * kLCH];
    __shared__ float wl[COUT * CIN * kTaps];
    const int tiles_x = (a.nx + kTX - 1) / kTX;
    const int tx0 = (blockIdx.x % tiles_x) * kTX, ty0 = (blockIdx.x / tiles_x) * kTY;
    const int z0 = blockIdx.y * kTZ;
    const int b = blockIdx.z;
    for (int i = threadIdx.x; i < COUT * CIN * kTaps; i += kConvThreads) wl[i] = (float)w[i];
    stage_tile<CIN, IN_MODE>(a, in + (int64_t)b * a.in_bs, tile, tx0, ty0, z0);
    __syncthreads();
    const int x = threadIdx.x % kTX, y = threadIdx.x / kTX;
    float acc[kTZ][COUT];
#pragma unroll
    for (int z = 0; z < kTZ; ++z)
#pragma unroll
        for (int c = 0; c < COUT; ++c) acc[z][c] = 0.f;
    for (int ci = 0; ci < CIN; ++ci) {
        const float* tc = tile + ci * kLCH;
#pragma unroll 1
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
        for (int t9 = 0; t9 < 9; ++t9) {
            const int t = dz * 9 + t9, dy = t9 / 3, dx = t9 % 3;
            float wv[COUT];
#pragma unroll
            for (int c = 0; c < COUT; ++c) wv[c] = wl[(c * CIN + ci) * kTaps + t];
#pragma unroll
            for (int z = 0; z < kTZ; ++z) {
                const float xv = tc[((z + dz) * kLY + (y + dy)) * kLX + x + dx];
#pragma unroll
                for (int c = 0; c < COUT; ++c) acc[z][c] += wv[c] * xv;
            }
        }
    }
    const int gx = tx0 + x, gy = ty0 + y;
    if (gx >= a.nx || gy >= a.ny) return;
#pragma unroll
    for (int z = 0; z < kTZ; ++z) {
        const int gz = z0 + z;
        if (gz >= a.nz) break;
        const int64_t gi = gx + (int64_t)a.nx * (gy + (int64_t)a.ny * gz);
        if constexpr (OUT_MODE == OUT_UV) {
            out[(int64_t)b * a.out_bs + gi] = acc[z][0] + a.v[gi] * acc[z][1];
        } else {
            const float vv = (OUT_MODE == OUT_GATE) ? a.v[gi] : 0.f;
#pragma unroll
            for (int c = 0; c < COUT; ++c) {
                float val = acc[z][c];
                if constexpr (OUT_MODE == OUT_GATE) {
                    if (pre) pre[(int64_t)b * a.pre_bs + c * a.ng + gi] = val;
                    val *= 1.f + tanhf(a.ga[c] * vv + a.gb[c]);
                }
                out[(int64_t)b * a.out_bs + c * a.ng + gi] = val;
            }
        }
    }
}

// ------------------------------------------------------------ MFMA convolution ----
// Inference form of conv3_kernel on the matrix cores (BF16 activations and weights, FP32
// accumulation): one wavefront computes 32 x-points x 4 z-planes of one y-row for up to 8
// output channels as v_mfma_f32_32x32x16_bf16 products, summed over the 9 (dy, dx) taps:
//   A (32 points x K),  K = (input plane zin of 6) x (input channel of 8),
//   B (K x 32),         column n = co * 4 + zo, B[(zin, ci)][n] = w[co][ci][zin - zo][dy][dx]
//                       (zero outside the 3-tap band),
// so 27 MFMAs per wavefront and tile.  The input tile (6 planes x 6 rows x 34 columns x 8
// channels, BF16, channel-fastest so a lane's 8 K-values are one 16-byte LDS read) is staged
// with the periodic wrap; the B fragments of all 27 (tap, K-step) pairs stay in registers.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int kMX = 32, kMY = 4, kMZ = 4;
constexpr int kMLX = kMX + 2, kMLY = kMY + 2, kMLZ = kMZ + 2;
constexpr int kMThreads = 64 * kMY;

__device__ __forceinline__ int modi(int i, int n) { return ((i % n) + n) % n; }

template <int CIN, int COUT, int IN_MODE, int OUT_MODE>
__global__ __launch_bounds__(kMThreads) void conv3_mfma_kernel(ConvArgs a, const float* __restrict__ w,
                                                               const float* __restrict__ in, float* __restrict__ out) {
    static_assert(CIN <= 8 && COUT <= 8, "at most 8 channels");
    __shared__ __align__(16) __bf16 tile[kMLZ * kMLY * kMLX * 8];
    const int tiles_x = (a.nx + kMX - 1) / kMX;
    const int x0 = (blockIdx.x % tiles_x) * kMX, y0 = (blockIdx.x / tiles_x) * kMY;
    const int z0 = blockIdx.y * kMZ;
    const int b = blockIdx.z;
    const float* inb = in + (int64_t)b * a.in_bs;
    for (int i = threadIdx.x; i < kMLZ * kMLY * kMLX; i += kMThreads) {
        const int xx = i % kMLX, yy = (i / kMLX) % kMLY, zz = i / (kMLX * kMLY);
        const int gx = modi(x0 + xx - 1, a.nx), gy = modi(y0 + yy - 1, a.ny), gz = modi(z0 + zz - 1, a.nz);
        const int64_t gi = gx + (int64_t)a.nx * (gy + (int64_t)a.ny * gz);
        float vals[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) vals[c] = 0.f;
        if constexpr (IN_MODE == IN_UV) {
            const float u = inb[gi];
            vals[0] = u;
            vals[1] = a.v[gi] * u;
        } else {
#pragma unroll
            for (int c = 0; c < CIN; ++c) vals[c] = inb[c * a.ng + gi];
        }
        bf16x8_t pk;
#pragma unroll
        for (int c = 0; c < 8; ++c) pk[c] = (__bf16)vals[c];
        *reinterpret_cast<bf16x8_t*>(&tile[i * 8]) = pk;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 31, kg = lane >> 5;
    const int co = n >> 2, zo = n & 3;
    bf16x8_t bw[27];
#pragma unroll
    for (int sh = 0; sh < 9; ++sh) {
        const int dy = sh / 3, dx = sh % 3;
#pragma unroll
        for (int kb = 0; kb < 3; ++kb) {
            const int dz = 2 * kb + kg - zo;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                float val = 0.f;
                if (co < COUT && j < CIN && dz >= 0 && dz <= 2) val = w[(co * CIN + j) * kTaps + dz * 9 + dy * 3 + dx];
                bw[sh * 3 + kb][j] = (__bf16)val;
            }
        }
    }
    __syncthreads();
    f32x16_t acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    const int r = lane & 31;
#pragma unroll
    for (int sh = 0; sh < 9; ++sh) {
        const int dy = sh / 3, dx = sh % 3;
#pragma unroll
        for (int kb = 0; kb < 3; ++kb) {
            const int zin = 2 * kb + kg;
            const bf16x8_t av =
                *reinterpret_cast<const bf16x8_t*>(&tile[((zin * kMLY + (wave + dy)) * kMLX + (r + dx)) * 8]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bw[sh * 3 + kb], acc, 0, 0, 0);
        }
    }
    const int gy = y0 + wave;
    const int gz = z0 + zo;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int gx = x0 + 8 * (i / 4) + 4 * kg + (i % 4);
        float partner = 0.f;
        if constexpr (OUT_MODE == OUT_UV) partner = __shfl(acc[i], lane + 4, 64);
        if (gx >= a.nx || gy >= a.ny || gz >= a.nz) continue;
        const int64_t gi = gx + (int64_t)a.nx * (gy + (int64_t)a.ny * gz);
        if constexpr (OUT_MODE == OUT_UV) {
            if (co == 0) out[(int64_t)b * a.out_bs + gi] = acc[i] + a.v[gi] * partner;
        } else {
            if (co < COUT) {
                float val = acc[i];
                if constexpr (OUT_MODE == OUT_GATE) val *= 1.f + tanhf(a.ga[co] * a.v[gi] + a.gb[co]);
                out[(int64_t)b * a.out_bs + co * a.ng + gi] = val;
            }
        }
    }
}

// partial[wg][k] = sum over the tile of delta[b][co](p) x[b][ci](p + d_tap), k = (co CIN + ci) 27 + tap
template <int CIN, int COUT, int IN_MODE>
__global__ __launch_bounds__(kConvThreads) void conv3_wgrad_kernel(ConvArgs a, const float* __restrict__ in,
                                                                   const float* __restrict__ delta, int64_t delta_bs,
                                                                   float* __restrict__ partial) {
    __shared__ float tile[CIN * kLCH];
    __shared__ float dt[COUT * kTZ * kTY * kTX];
    const int tiles_x = (a.nx + kTX - 1) / kTX;
    const int tx0 = (blockIdx.x % tiles_x) * kTX, ty0 = (blockIdx.x / tiles_x) * kTY;
    const int z0 = blockIdx.y * kTZ;
    const int b = blockIdx.z;
    stage_tile<CIN, IN_MODE>(a, in + (int64_t)b * a.in_bs, tile, tx0, ty0, z0);
    const float* db = delta + (int64_t)b * delta_bs;
    for (int i = threadIdx.x; i < COUT * kTZ * kTY * kTX; i += kConvThreads) {
        const int x = i % kTX, y = (i / kTX) % kTY, z = (i / (kTX * kTY)) % kTZ, c = i / (kTX * kTY * kTZ);
        const int gx = tx0 + x, gy = ty0 + y, gz = z0 + z;
        float d = 0.f;
        if (gx < a.nx && gy < a.ny && gz < a.nz) d = db[c * a.ng + gx + (int64_t)a.nx * (gy + (int64_t)a.ny * gz)];
        dt[i] = d;
    }
    __syncthreads();
    const int64_t wg = blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
    constexpr int NW = COUT * CIN * kTaps;
    for (int k = threadIdx.x; k < NW; k += kConvThreads) {
        const int co = k / (CIN * kTaps), ci = (k / kTaps) % CIN, t = k % kTaps;
        const int dz = t / 9, dy = (t / 3) % 3, dx = t % 3;
        const float* tc = tile + ci * kLCH;
        const float* dc = dt + co * kTZ * kTY * kTX;
        float s = 0.f;
        for (int z = 0; z < kTZ; ++z)
            for (int y = 0; y < kTY; ++y) {
                const float* trow = tc + ((z + dz) * kLY + (y + dy)) * kLX + dx;
                const float* drow = dc + (z * kTY + y) * kTX;
#pragma unroll 4
                for (int x = 0; x < kTX; ++x) s += drow[x] * trow[x];
            }
        partial[wg * NW + k] = s;
    }
}

// Gate parameter gradients: dA_c = sum dh_c a_c (1 - t^2) v, dB_c = sum dh_c a_c (1 - t^2),
// t = tanh(A_c v + B_c); per-workgroup partials [wg][2 C].
__global__ __launch_bounds__(256) void gate_grad_kernel(int64_t ng, int nb, const float* __restrict__ dh,
                                                        const float* __restrict__ act, const float* __restrict__ v,
                                                        const float* __restrict__ ga, const float* __restrict__ gb,
                                                        float* __restrict__ partial) {
    __shared__ float red[256 / 64][2 * kC];
    float sa[kC], sb[kC];
#pragma unroll
    for (int c = 0; c < kC; ++c) sa[c] = sb[c] = 0.f;
    const int64_t total = ng * nb;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = i / ng, p = i - b * ng;
        const float vv = v[p];
#pragma unroll
        for (int c = 0; c < kC; ++c) {
            const int64_t o = (b * kC + c) * ng + p;
            const float t = tanhf(ga[c] * vv + gb[c]);
            const float q = dh[o] * act[o] * (1.f - t * t);
            sa[c] += q * vv;
            sb[c] += q;
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < kC; ++c) {
        float x = sa[c], y = sb[c];
        for (int s = 32; s >= 1; s >>= 1) {
            x += __shfl_xor(x, s, 64);
            y += __shfl_xor(y, s, 64);
        }
        if (lane == 0) {
            red[wave][c] = x;
            red[wave][kC + c] = y;
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * kC) {
        float s = 0.f;
        for (int w = 0; w < 256 / 64; ++w) s += red[w][threadIdx.x];
        partial[blockIdx.x * 2 * kC + threadIdx.x] = s;
    }
}

// grad[k] (+)= sum_w partial[w][k], FP64 accumulation
__global__ void reduce_partials_kernel(const float* __restrict__ partial, int64_t nwg, int nw, float* __restrict__ grad,
                                       int accumulate) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nw) return;
    double s = 0.0;
    for (int64_t w = 0; w < nwg; ++w) s += (double)partial[w * nw + k];
    grad[k] = (float)(accumulate ? grad[k] + s : s);
}

// W^T of one layer: wt[ci][co][t] = w[co][ci][26 - t]
__global__ void transpose_weights_kernel(const float* __restrict__ w, float* __restrict__ wt, int cout, int cin) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cout * cin * kTaps) return;
    const int co = k / (cin * kTaps), ci = (k / kTaps) % cin, t = k % kTaps;
    wt[(ci * cout + co) * kTaps + (kTaps - 1 - t)] = w[k];
}

__global__ void to_bf16_kernel(const float* __restrict__ w, __bf16* __restrict__ wb, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) wb[k] = (__bf16)w[k];
}

// e = y - t; loss partial sum e^2; g = 2 e / z  (in place over y)
__global__ __launch_bounds__(256) void loss_kernel(float* __restrict__ y, const float* __restrict__ t, int64_t n,
                                                   double inv_z, double* __restrict__ partial) {
    __shared__ double red[4];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float e = y[i] - t[i];
        s += (double)e * e;
        y[i] = (float)(2.0 * inv_z * e);
    }
    for (int sh = 32; sh >= 1; sh >>= 1) s += __shfl_xor(s, sh, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2, float c1, float c2, float eps) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

template <class Ts, class Td>
__global__ void cnn_convert_kernel(const Ts* __restrict__ a, Td* __restrict__ b, int64_t n, double c) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (Td)((double)a[i] * c);
}

// y += x (FP32 into FP32 or FP64)
template <class T>
__global__ void add_kernel(T* __restrict__ y, const float* __restrict__ x, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (T)(y[i] + (T)x[i]);
}

static int grid1(int64_t n) { return (int)std::min<int64_t>(4096, (n + 255) / 256); }

}  // namespace qg

// ------------------------------------------------------------------------ handle ----
struct qg_cnn {
    qg_ham* h = nullptr;
    double shift = 1.0;                 // s of K = (T + s)^-1
    std::vector<float> params;          // host copy (FP32 master weights)
    qg::DevBuf<float> p, grad, m, v2;   // parameters, gradient, Adam moments
    qg::DevBuf<float> wt;               // transposed weights of the three layers
    qg::DevBuf<__bf16> pb;              // BF16 inference weights
    qg::DevBuf<float> wtb;              // transposed BF16-rounded weights (as FP32)
    qg::DevBuf<float> pbf;              // BF16-rounded forward weights (as FP32)
    qg::DevBuf<float> vnorm;            // (V - mean) / std
    bool dirty = true;                  // derived weight copies out of date
    bool mfma = false;                  // preconditioner network on the matrix cores
    int64_t adam_t = 0;
    // activation workspace (chunk of bands)
    qg::DevBuf<float> ws;
    qg::DevBuf<float> partial;
    qg::DevBuf<double> dpart;
    // training pairs
    qg::DevBuf<float> U, Tt;
    std::vector<double> zn;
    int npairs = 0;
};

namespace qg {

static ConvArgs conv_args(const qg_cnn* c) {
    const qg_grid& g = c->h->grid;
    ConvArgs a{};
    a.nx = g.n[0]; a.ny = g.n[1]; a.nz = g.n[2];
    a.ng = g.ng();
    a.v = c->vnorm.p;
    return a;
}

static dim3 conv_grid(const qg_grid& g, int nb) {
    return dim3((unsigned)(((g.n[0] + kTX - 1) / kTX) * ((g.n[1] + kTY - 1) / kTY)), (unsigned)((g.n[2] + kTZ - 1) / kTZ),
                (unsigned)nb);
}

// weight views (FP32 device pointers) of one parameter set
struct Weights {
    const float* w1; const float* w2; const float* w3;      // forward
    const float* t1; const float* t2; const float* t3;      // transposed (adjoint)
    const float* a1; const float* b1; const float* a2; const float* b2;
};

static void refresh_derived(qg_cnn* c, hipStream_t s) {
    if (!c->dirty) return;
    c->wt.alloc(kOffA1);
    c->pbf.alloc(kNumParams);
    c->wtb.alloc(kOffA1);
    c->pb.alloc(kNumParams);
    // BF16 copy and its FP32 image (the inference weights)
    hipLaunchKernelGGL(to_bf16_kernel, dim3((kNumParams + 255) / 256), dim3(256), 0, s, c->p.p, c->pb.p, kNumParams);
    hipLaunchKernelGGL((cnn_convert_kernel<__bf16, float>), dim3(16), dim3(256), 0, s, c->pb.p, c->pbf.p,
                       (int64_t)kNumParams, 1.0);
    auto tr = [&](const float* src, float* dst) {
        hipLaunchKernelGGL(transpose_weights_kernel, dim3(8), dim3(256), 0, s, src + kOffW1, dst + kOffW1, kC, 2);
        hipLaunchKernelGGL(transpose_weights_kernel, dim3(8), dim3(256), 0, s, src + kOffW2, dst + kOffW2, kC, kC);
        hipLaunchKernelGGL(transpose_weights_kernel, dim3(8), dim3(256), 0, s, src + kOffW3, dst + kOffW3, 1, kC);
    };
    tr(c->p.p, c->wt.p);
    tr(c->pbf.p, c->wtb.p);
    QG_HIP(hipGetLastError());
    c->dirty = false;
}

static Weights weights(qg_cnn* c, bool bf16) {
    const float* f = bf16 ? c->pbf.p : c->p.p;
    const float* t = bf16 ? c->wtb.p : c->wt.p;
    return Weights{f + kOffW1, f + kOffW2, f + kOffW3, t + kOffW1, t + kOffW2, t + kOffW3,
                   f + kOffA1, f + kOffB1, f + kOffA2, f + kOffB2};
}

// Activations of one forward pass over nb bands: a1, h1, a2, h2 (8 channels each).
struct Acts {
    float* a1; float* h1; float* a2; float* h2;
};

// w = N u (u: nb x ng); activations stored when acts != null (else scratch from s8a/s8b)
static void net_forward(qg_cnn* c, const Weights& W, int nb, const float* u, float* w, const Acts* acts, float* s8a,
                        float* s8b, hipStream_t s) {
    const qg_grid& g = c->h->grid;
    const int64_t ng = g.ng();
    ConvArgs a = conv_args(c);
    const dim3 grid = conv_grid(g, nb);
    float* h1 = acts ? acts->h1 : s8a;
    float* h2 = acts ? acts->h2 : s8b;
    a.in_bs = ng; a.out_bs = kC * ng; a.pre_bs = kC * ng; a.ga = W.a1; a.gb = W.b1;
    hipLaunchKernelGGL((conv3_kernel<2, kC, IN_UV, OUT_GATE, float>), grid, dim3(kConvThreads), 0, s, a, W.w1, u, h1,
                       acts ? acts->a1 : nullptr);
    a.in_bs = kC * ng; a.ga = W.a2; a.gb = W.b2;
    hipLaunchKernelGGL((conv3_kernel<kC, kC, IN_PLAIN, OUT_GATE, float>), grid, dim3(kConvThreads), 0, s, a, W.w2, h1,
                       h2, acts ? acts->a2 : nullptr);
    a.out_bs = ng;
    hipLaunchKernelGGL((conv3_kernel<kC, 1, IN_PLAIN, OUT_PLAIN, float>), grid, dim3(kConvThreads), 0, s, a, W.w3, h2, w,
                       nullptr);
    QG_HIP(hipGetLastError());
}

static dim3 mfma_grid(const qg_grid& g, int nb) {
    return dim3((unsigned)(((g.n[0] + kMX - 1) / kMX) * ((g.n[1] + kMY - 1) / kMY)), (unsigned)((g.n[2] + kMZ - 1) / kMZ),
                (unsigned)nb);
}

// w = N u and q = N^T y on the matrix cores (BF16 weights and activations), scratch s8a/s8b
static void net_forward_mfma(qg_cnn* c, const Weights& W, int nb, const float* u, float* w, float* s8a, float* s8b,
                             hipStream_t s) {
    const qg_grid& g = c->h->grid;
    const int64_t ng = g.ng();
    ConvArgs a = conv_args(c);
    const dim3 grid = mfma_grid(g, nb);
    a.in_bs = ng; a.out_bs = kC * ng; a.ga = W.a1; a.gb = W.b1;
    hipLaunchKernelGGL((conv3_mfma_kernel<2, kC, IN_UV, OUT_GATE>), grid, dim3(kMThreads), 0, s, a, W.w1, u, s8a);
    a.in_bs = kC * ng; a.ga = W.a2; a.gb = W.b2;
    hipLaunchKernelGGL((conv3_mfma_kernel<kC, kC, IN_PLAIN, OUT_GATE>), grid, dim3(kMThreads), 0, s, a, W.w2, s8a, s8b);
    a.out_bs = ng;
    hipLaunchKernelGGL((conv3_mfma_kernel<kC, 1, IN_PLAIN, OUT_PLAIN>), grid, dim3(kMThreads), 0, s, a, W.w3, s8b, w);
    QG_HIP(hipGetLastError());
}

static void net_adjoint_mfma(qg_cnn* c, const Weights& W, int nb, const float* y, float* q, float* s8a, float* s8b,
                             hipStream_t s) {
    const qg_grid& g = c->h->grid;
    const int64_t ng = g.ng();
    ConvArgs a = conv_args(c);
    const dim3 grid = mfma_grid(g, nb);
    a.in_bs = ng; a.out_bs = kC * ng; a.ga = W.a2; a.gb = W.b2;
    hipLaunchKernelGGL((conv3_mfma_kernel<1, kC, IN_PLAIN, OUT_GATE>), grid, dim3(kMThreads), 0, s, a, W.t3, y, s8a);
    a.in_bs = kC * ng; a.ga = W.a1; a.gb = W.b1;
    hipLaunchKernelGGL((conv3_mfma_kernel<kC, kC, IN_PLAIN, OUT_GATE>), grid, dim3(kMThreads), 0, s, a, W.t2, s8a, s8b);
    a.out_bs = ng;
    hipLaunchKernelGGL((conv3_mfma_kernel<kC, 2, IN_PLAIN, OUT_UV>), grid, dim3(kMThreads), 0, s, a, W.t1, s8b, q);
    QG_HIP(hipGetLastError());
}

// q = N^T y; when dh2/dh1 are given, the pre-gate adjoint activations are stored there and the
// post-gate ones in da2/da1 (else scratch)
static void net_adjoint(qg_cnn* c, const Weights& W, int nb, const float* y, float* q, float* da2, float* da1,
                        float* dh2, float* dh1, hipStream_t s) {
    const qg_grid& g = c->h->grid;
    const int64_t ng = g.ng();
    ConvArgs a = conv_args(c);
    const dim3 grid = conv_grid(g, nb);
    a.in_bs = ng; a.out_bs = kC * ng; a.pre_bs = kC * ng; a.ga = W.a2; a.gb = W.b2;
    hipLaunchKernelGGL((conv3_kernel<1, kC, IN_PLAIN, OUT_GATE, float>), grid, dim3(kConvThreads), 0, s, a, W.t3, y, da2,
                       dh2);
    a.in_bs = kC * ng; a.ga = W.a1; a.gb = W.b1;
    hipLaunchKernelGGL((conv3_kernel<kC, kC, IN_PLAIN, OUT_GATE, float>), grid, dim3(kConvThreads), 0, s, a, W.t2, da2,
                       da1, dh1);
    a.out_bs = ng;
    hipLaunchKernelGGL((conv3_kernel<kC, 2, IN_PLAIN, OUT_UV, float>), grid, dim3(kConvThreads), 0, s, a, W.t1, da1, q,
                       nullptr);
    QG_HIP(hipGetLastError());
}

// Weight and gate gradients of <upstream, N x> over nb bands for a forward pass with stored
// activations (input x, acts); accumulated into c->grad.  Scratch: da2, da1, dh2, dh1 (8 ch).
static void net_backward(qg_cnn* c, const Weights& W, int nb, const float* x, const Acts& acts, const float* upstream,
                         float* da2, float* da1, float* dh2, float* dh1, hipStream_t s) {
    const qg_grid& g = c->h->grid;
    const int64_t ng = g.ng();
    ConvArgs a = conv_args(c);
    const dim3 grid = conv_grid(g, nb);
    const int64_t nwg = (int64_t)grid.x * grid.y * grid.z;
    const int ggrid = 512;
    c->partial.alloc((size_t)std::max<int64_t>(nwg * kC * kC * kTaps, ggrid * 2 * kC));
    float* part = c->partial.p;
    // dW3 = corr(upstream, h2)
    a.in_bs = kC * ng;
    hipLaunchKernelGGL((conv3_wgrad_kernel<kC, 1, IN_PLAIN>), grid, dim3(kConvThreads), 0, s, a, acts.h2, upstream, ng,
                       part);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((kC * kTaps + 255) / 256), dim3(256), 0, s, part, nwg, kC * kTaps,
                       c->grad.p + kOffW3, 1);
    // dh2 = W3^T upstream, da2 = dh2 g2
    a.in_bs = ng; a.out_bs = kC * ng; a.pre_bs = kC * ng; a.ga = W.a2; a.gb = W.b2;
    hipLaunchKernelGGL((conv3_kernel<1, kC, IN_PLAIN, OUT_GATE, float>), grid, dim3(kConvThreads), 0, s, a, W.t3,
                       upstream, da2, dh2);
    hipLaunchKernelGGL(gate_grad_kernel, dim3(ggrid), dim3(256), 0, s, ng, nb, dh2, acts.a2, c->vnorm.p, W.a2, W.b2,
                       part);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(64), 0, s, part, (int64_t)ggrid, 2 * kC,
                       c->grad.p + kOffA2, 1);
    // dW2 = corr(da2, h1)
    a.in_bs = kC * ng;
    hipLaunchKernelGGL((conv3_wgrad_kernel<kC, kC, IN_PLAIN>), grid, dim3(kConvThreads), 0, s, a, acts.h1, da2, kC * ng,
                       part);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((kC * kC * kTaps + 255) / 256), dim3(256), 0, s, part, nwg,
                       kC * kC * kTaps, c->grad.p + kOffW2, 1);
    // dh1 = W2^T da2, da1 = dh1 g1
    a.in_bs = kC * ng; a.out_bs = kC * ng; a.ga = W.a1; a.gb = W.b1;
    hipLaunchKernelGGL((conv3_kernel<kC, kC, IN_PLAIN, OUT_GATE, float>), grid, dim3(kConvThreads), 0, s, a, W.t2, da2,
                       da1, dh1);
    hipLaunchKernelGGL(gate_grad_kernel, dim3(ggrid), dim3(256), 0, s, ng, nb, dh1, acts.a1, c->vnorm.p, W.a1, W.b1,
                       part);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(64), 0, s, part, (int64_t)ggrid, 2 * kC,
                       c->grad.p + kOffA1, 1);
    // dW1 = corr(da1, [x, v x])
    a.in_bs = ng;
    hipLaunchKernelGGL((conv3_wgrad_kernel<2, kC, IN_UV>), grid, dim3(kConvThreads), 0, s, a, x, da1, kC * ng, part);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((kC * 2 * kTaps + 255) / 256), dim3(256), 0, s, part, nwg,
                       kC * 2 * kTaps, c->grad.p + kOffW1, 1);
    QG_HIP(hipGetLastError());
}

// Workspace layout for a training batch of nb samples: 4 forward activation sets of 8 channels
// for each of the two passes, plus 4 adjoint scratch sets, plus single-channel vectors.
struct TrainWS {
    Acts au, ag;
    float *da2, *da1, *dh2, *dh1;
    float *w_u, *q, *g, *w_g;
};

static TrainWS train_ws(qg_cnn* c, int nb) {
    const int64_t ng = c->h->grid.ng();
    const int64_t c8 = (int64_t)nb * kC * ng, c1 = (int64_t)nb * ng;
    c->ws.alloc((size_t)(12 * c8 + 4 * c1));
    float* p = c->ws.p;
    TrainWS t{};
    t.au = Acts{p, p + c8, p + 2 * c8, p + 3 * c8};
    p += 4 * c8;
    t.ag = Acts{p, p + c8, p + 2 * c8, p + 3 * c8};
    p += 4 * c8;
    t.da2 = p; t.da1 = p + c8; t.dh2 = p + 2 * c8; t.dh1 = p + 3 * c8;
    p += 4 * c8;
    t.w_u = p; t.q = p + c1; t.g = p + 2 * c1; t.w_g = p + 3 * c1;
    return t;
}

// loss and gradient (into c->grad) of one batch: u, t (nb x ng, FP32), znorm = sum |z*|^2
static double loss_grad(qg_cnn* c, int nb, const float* u, const float* t, double znorm, hipStream_t s) {
    refresh_derived(c, s);
    const int64_t ng = c->h->grid.ng();
    const Weights W = weights(c, false);
    TrainWS ws = train_ws(c, nb);
    QG_HIP(hipMemsetAsync(c->grad.p, 0, kNumParams * sizeof(float), s));
    // w = N u (activations kept), q = N^T w, y = K q
    net_forward(c, W, nb, u, ws.w_u, &ws.au, nullptr, nullptr, s);
    net_adjoint(c, W, nb, ws.w_u, ws.q, ws.da2, ws.da1, nullptr, nullptr, s);
    precond_apply(c->h, QG_PRECOND_KINETIC, c->shift, QG_F32, nb, ws.q, ng, s);
    // e = y - t, loss, g = 2 K e / Z
    const int lg = 256;
    c->dpart.alloc(lg);
    hipLaunchKernelGGL(loss_kernel, dim3(lg), dim3(256), 0, s, ws.q, t, (int64_t)nb * ng, 1.0 / znorm, c->dpart.p);
    std::vector<double> lp(lg);
    QG_HIP(hipMemcpyAsync(lp.data(), c->dpart.p, lg * sizeof(double), hipMemcpyDeviceToHost, s));
    precond_apply(c->h, QG_PRECOND_KINETIC, c->shift, QG_F32, nb, ws.q, ng, s);
    float* gvec = ws.q;
    // N g with activations
    net_forward(c, W, nb, gvec, ws.w_g, &ws.ag, nullptr, nullptr, s);
    // <w_u, dN g> + <N g, dN u>
    net_backward(c, W, nb, gvec, ws.ag, ws.w_u, ws.da2, ws.da1, ws.dh2, ws.dh1, s);
    net_backward(c, W, nb, u, ws.au, ws.w_g, ws.da2, ws.da1, ws.dh2, ws.dh1, s);
    QG_HIP(hipStreamSynchronize(s));
    double loss = 0.0;
    for (double v : lp) loss += v;
    return loss / znorm;
}

// M r in place for nb bands (FP32 or FP64 block, ld = Ng), chunked over bands.
void cnn_precond_apply(qg_cnn* c, int dtype, int nbands, void* r, int64_t ld, hipStream_t s) {
    const qg_grid& g = c->h->grid;
    const int64_t ng = g.ng();
    QG_REQUIRE(ld == ng, "learned preconditioner expects ld == Ng");
    refresh_derived(c, s);
    const Weights W = weights(c, true);
    // per band: u (1), two 8-channel sets, w (1), y (1)
    const int64_t per = (2 * kC + 3) * ng;
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(nbands, (1ll << 30) / (per * 4)));
    c->ws.alloc((size_t)chunk * per);
    float* u = c->ws.p;
    float* s8a = u + (int64_t)chunk * ng;
    float* s8b = s8a + (int64_t)chunk * kC * ng;
    float* w = s8b + (int64_t)chunk * kC * ng;
    float* y = w + (int64_t)chunk * ng;
    // u = K r in place
    precond_apply(c->h, QG_PRECOND_KINETIC, c->shift, dtype, nbands, r, ld, s);
    for (int b0 = 0; b0 < nbands; b0 += chunk) {
        const int nb = std::min(chunk, nbands - b0);
        const int64_t n = (int64_t)nb * ng;
        if (dtype == QG_F32) {
            QG_HIP(hipMemcpyAsync(u, (float*)r + (int64_t)b0 * ng, n * 4, hipMemcpyDeviceToDevice, s));
        } else {
            hipLaunchKernelGGL((cnn_convert_kernel<double, float>), dim3(grid1(n)), dim3(256), 0, s,
                               (const double*)r + (int64_t)b0 * ng, u, n, 1.0);
        }
        if (c->mfma) {
            net_forward_mfma(c, W, nb, u, w, s8a, s8b, s);
            net_adjoint_mfma(c, W, nb, w, y, s8a, s8b, s);
        } else {
            net_forward(c, W, nb, u, w, nullptr, s8a, s8b, s);
            net_adjoint(c, W, nb, w, y, s8a, s8b, nullptr, nullptr, s);
        }
        precond_apply(c->h, QG_PRECOND_KINETIC, c->shift, QG_F32, nb, y, ng, s);
        if (dtype == QG_F32)
            hipLaunchKernelGGL((add_kernel<float>), dim3(grid1(n)), dim3(256), 0, s, (float*)r + (int64_t)b0 * ng, y, n);
        else
            hipLaunchKernelGGL((add_kernel<double>), dim3(grid1(n)), dim3(256), 0, s, (double*)r + (int64_t)b0 * ng, y,
                               n);
    }
    QG_HIP(hipGetLastError());
}

void cnn_free(qg_cnn* c) { delete c; }

}  // namespace qg

namespace qg {

// ------------------------------------------------------------- training pairs ----
// per-band dot products out[k] = <a_k, b_k> (FP64), one workgroup per band
__global__ __launch_bounds__(256) void bdot_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                   int64_t n, double* __restrict__ out) {
    __shared__ double red[4];
    const double* pa = a + (int64_t)blockIdx.x * n;
    const double* pb = b + (int64_t)blockIdx.x * n;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += pa[i] * pb[i];
    for (int sh = 32; sh >= 1; sh >>= 1) s += __shfl_xor(s, sh, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// y_k = alpha_k x_k + beta_k y_k  (per-band coefficient arrays, either may be null = 1 / 0)
__global__ void baxpby_kernel(double* __restrict__ y, const double* __restrict__ x, const double* __restrict__ alpha,
                              double alpha_sign, const double* __restrict__ beta, int64_t n, int nb) {
    const int64_t total = n * nb;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = i / n;
        const double a = alpha ? alpha_sign * alpha[k] : alpha_sign;
        const double b = beta ? beta[k] : 1.0;
        y[i] = a * x[i] + b * y[i];
    }
}

// r_k = H x_k - e_k x_k
__global__ void resid_kernel(double* __restrict__ r, const double* __restrict__ x, const double* __restrict__ e,
                             int64_t n, int nb) {
    const int64_t total = n * nb;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        r[i] -= e[i / n] * x[i];
}

// (z - u) / rms and u / rms into the FP32 pair store; zn_k = |z_k / rms_k|^2
__global__ void pair_store_kernel(const double* __restrict__ u, const double* __restrict__ z,
                                  const double* __restrict__ inv_rms, float* __restrict__ U, float* __restrict__ Tt,
                                  int64_t n, int nb) {
    const int64_t total = n * nb;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const double c = inv_rms[i / n];
        U[i] = (float)(u[i] * c);
        Tt[i] = (float)((z[i] - u[i]) * c);
    }
}

// Solve (T + V - sigma) z = r for nb bands (FP64, ld = ng) by preconditioned CG with K.
static int solve_local(qg_cnn* c, double sigma, int nb, const double* r, double* z, double tol, int maxit,
                       hipStream_t s) {
    qg_ham* h = c->h;
    const int64_t ng = h->grid.ng();
    const int64_t n = (int64_t)nb * ng;
    DevBuf<double> rr, zz, p, ap, d;
    rr.alloc(n); zz.alloc(n); p.alloc(n); ap.alloc(n); d.alloc(4 * (size_t)nb);
    double* rz = d.p;
    double* pap = d.p + nb;
    double* alpha = d.p + 2 * nb;
    double* beta = d.p + 3 * nb;
    const int gs = grid1(n);
    QG_HIP(hipMemsetAsync(z, 0, n * 8, s));
    QG_HIP(hipMemcpyAsync(rr.p, r, n * 8, hipMemcpyDeviceToDevice, s));
    QG_HIP(hipMemcpyAsync(zz.p, r, n * 8, hipMemcpyDeviceToDevice, s));
    precond_apply(h, QG_PRECOND_KINETIC, c->shift, QG_F64, nb, zz.p, ng, s);
    QG_HIP(hipMemcpyAsync(p.p, zz.p, n * 8, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(bdot_kernel, dim3(nb), dim3(256), 0, s, rr.p, zz.p, ng, rz);
    std::vector<double> h_rz(nb), h_pap(nb), h_rr(nb), r0(nb), h_al(nb), h_be(nb);
    hipLaunchKernelGGL(bdot_kernel, dim3(nb), dim3(256), 0, s, r, r, ng, pap);
    QG_HIP(hipMemcpyAsync(r0.data(), pap, nb * 8, hipMemcpyDeviceToHost, s));
    QG_HIP(hipMemcpyAsync(h_rz.data(), rz, nb * 8, hipMemcpyDeviceToHost, s));
    QG_HIP(hipStreamSynchronize(s));
    int it = 0;
    for (it = 0; it < maxit; ++it) {
        local_hamiltonian_apply(h, QG_F64, nb, p.p, ng, ap.p, ng, s);
        hipLaunchKernelGGL(baxpby_kernel, dim3(gs), dim3(256), 0, s, ap.p, p.p, (const double*)nullptr, -sigma,
                           (const double*)nullptr, ng, nb);
        hipLaunchKernelGGL(bdot_kernel, dim3(nb), dim3(256), 0, s, p.p, ap.p, ng, pap);
        QG_HIP(hipMemcpyAsync(h_pap.data(), pap, nb * 8, hipMemcpyDeviceToHost, s));
        QG_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < nb; ++k) {
            QG_REQUIRE(h_pap[k] > 0.0, "local operator is not positive definite at this sigma");
            h_al[k] = h_rz[k] / h_pap[k];
        }
        QG_HIP(hipMemcpyAsync(alpha, h_al.data(), nb * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(baxpby_kernel, dim3(gs), dim3(256), 0, s, z, p.p, alpha, 1.0, (const double*)nullptr, ng, nb);
        hipLaunchKernelGGL(baxpby_kernel, dim3(gs), dim3(256), 0, s, rr.p, ap.p, alpha, -1.0, (const double*)nullptr,
                           ng, nb);
        hipLaunchKernelGGL(bdot_kernel, dim3(nb), dim3(256), 0, s, rr.p, rr.p, ng, pap);
        QG_HIP(hipMemcpyAsync(h_rr.data(), pap, nb * 8, hipMemcpyDeviceToHost, s));
        QG_HIP(hipStreamSynchronize(s));
        bool done = true;
        for (int k = 0; k < nb; ++k) done = done && (h_rr[k] <= tol * tol * r0[k]);
        if (done) { ++it; break; }
        QG_HIP(hipMemcpyAsync(zz.p, rr.p, n * 8, hipMemcpyDeviceToDevice, s));
        precond_apply(h, QG_PRECOND_KINETIC, c->shift, QG_F64, nb, zz.p, ng, s);
        hipLaunchKernelGGL(bdot_kernel, dim3(nb), dim3(256), 0, s, rr.p, zz.p, ng, rz);
        std::vector<double> rz_new(nb);
        QG_HIP(hipMemcpyAsync(rz_new.data(), rz, nb * 8, hipMemcpyDeviceToHost, s));
        QG_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < nb; ++k) {
            h_be[k] = rz_new[k] / h_rz[k];
            h_rz[k] = rz_new[k];
        }
        QG_HIP(hipMemcpyAsync(beta, h_be.data(), nb * 8, hipMemcpyHostToDevice, s));
        // p = z + beta p
        hipLaunchKernelGGL(baxpby_kernel, dim3(gs), dim3(256), 0, s, p.p, zz.p, (const double*)nullptr, 1.0, beta, ng,
                           nb);
    }
    QG_HIP(hipStreamSynchronize(s));
    return it;
}

// Residual pairs from a block of solver iterates X (FP64, nb x ng) with Ritz values e.
static int collect_pairs(qg_cnn* c, int nb, const double* X, const double* e_host, double sigma, double tol,
                         int maxit, hipStream_t s) {
    qg_ham* h = c->h;
    const int64_t ng = h->grid.ng();
    const int64_t n = (int64_t)nb * ng;
    const int room = std::max(0, (int)(c->U.n / std::max<int64_t>(1, ng)) - c->npairs);
    const int take = std::min(nb, room);
    if (take <= 0) return 0;
    DevBuf<double> r, z, u, e, inv;
    r.alloc(n); z.alloc(n); u.alloc(n); e.alloc(nb); inv.alloc(nb);
    e.upload(e_host, nb, s);
    hamiltonian_apply(h, QG_F64, nb, X, ng, r.p, ng, s);
    hipLaunchKernelGGL(resid_kernel, dim3(grid1(n)), dim3(256), 0, s, r.p, X, e.p, ng, nb);
    const int iters = solve_local(c, sigma, take, r.p, z.p, tol, maxit, s);
    const int64_t nt = (int64_t)take * ng;
    QG_HIP(hipMemcpyAsync(u.p, r.p, nt * 8, hipMemcpyDeviceToDevice, s));
    precond_apply(h, QG_PRECOND_KINETIC, c->shift, QG_F64, take, u.p, ng, s);
    hipLaunchKernelGGL(bdot_kernel, dim3(take), dim3(256), 0, s, u.p, u.p, ng, inv.p);
    std::vector<double> uu(take), zz(take), ic(take);
    QG_HIP(hipMemcpyAsync(uu.data(), inv.p, take * 8, hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(bdot_kernel, dim3(take), dim3(256), 0, s, z.p, z.p, ng, e.p);
    QG_HIP(hipMemcpyAsync(zz.data(), e.p, take * 8, hipMemcpyDeviceToHost, s));
    QG_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < take; ++k) {
        const double rms2 = uu[k] / (double)ng;
        QG_REQUIRE(rms2 > 0.0, "zero residual in a training block");
        ic[k] = 1.0 / std::sqrt(rms2);
        c->zn.push_back(zz[k] / rms2);
    }
    inv.upload(ic.data(), take, s);
    hipLaunchKernelGGL(pair_store_kernel, dim3(grid1(nt)), dim3(256), 0, s, u.p, z.p, inv.p,
                       c->U.p + (int64_t)c->npairs * ng, c->Tt.p + (int64_t)c->npairs * ng, ng, take);
    QG_HIP(hipStreamSynchronize(s));
    c->npairs += take;
    return iters;
}

static void train(qg_cnn* c, int steps, int batch, double lr, uint64_t seed, double* first, double* last,
                  hipStream_t s) {
    QG_REQUIRE(c->npairs > 0, "no training pairs");
    const int64_t ng = c->h->grid.ng();
    batch = std::max(1, std::min(batch, c->npairs));
    DevBuf<float> bu, bt;
    bu.alloc((size_t)batch * ng);
    bt.alloc((size_t)batch * ng);
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int> pick(0, c->npairs - 1);
    const float b1 = 0.9f, b2 = 0.999f;
    double acc = 0.0;
    int nacc = 0;
    for (int it = 0; it < steps; ++it) {
        double znorm = 0.0;
        for (int k = 0; k < batch; ++k) {
            const int j = pick(rng);
            QG_HIP(hipMemcpyAsync(bu.p + (int64_t)k * ng, c->U.p + (int64_t)j * ng, ng * 4, hipMemcpyDeviceToDevice, s));
            QG_HIP(hipMemcpyAsync(bt.p + (int64_t)k * ng, c->Tt.p + (int64_t)j * ng, ng * 4, hipMemcpyDeviceToDevice, s));
            znorm += c->zn[j];
        }
        const double loss = loss_grad(c, batch, bu.p, bt.p, znorm, s);
        if (it == 0 && first) *first = loss;
        if (it >= steps - std::max(1, steps / 10)) {
            acc += loss;
            ++nacc;
        }
        ++c->adam_t;
        const float c1 = 1.f - std::pow(b1, (float)c->adam_t), c2 = 1.f - std::pow(b2, (float)c->adam_t);
        hipLaunchKernelGGL(adam_kernel, dim3((kNumParams + 255) / 256), dim3(256), 0, s, c->p.p, c->grad.p, c->m.p,
                           c->v2.p, kNumParams, (float)lr, b1, b2, c1, c2, 1e-8f);
        c->dirty = true;
    }
    QG_HIP(hipMemcpyAsync(c->params.data(), c->p.p, kNumParams * 4, hipMemcpyDeviceToHost, s));
    QG_HIP(hipStreamSynchronize(s));
    if (last) *last = nacc ? acc / nacc : 0.0;
}

}  // namespace qg

using namespace qg;

extern "C" {

int qg_cnn_num_params(void) { return kNumParams; }

int qg_cnn_create(qg_ham* h, double shift, int max_pairs, uint64_t seed, qg_cnn** out) {
    return guarded([&] {
        QG_REQUIRE(h && out, "null argument");
        QG_REQUIRE(shift > 0.0, "the kinetic shift must be positive");
        QG_REQUIRE(max_pairs >= 0, "bad pair capacity");
        auto c = std::make_unique<qg_cnn>();
        c->h = h;
        c->shift = shift;
        const char* mf = std::getenv("QG_CNN_MFMA");
        c->mfma = mf && *mf && std::atoi(mf) != 0;
        const int64_t ng = h->grid.ng();
        // normalised potential (from the FP64 copy)
        std::vector<double> V(ng);
        QG_HIP(hipMemcpy(V.data(), h->V64.p, ng * 8, hipMemcpyDeviceToHost));
        double mean = 0.0, var = 0.0;
        for (double x : V) mean += x;
        mean /= (double)ng;
        for (double x : V) var += (x - mean) * (x - mean);
        const double sd = std::sqrt(var / (double)ng);
        std::vector<float> vn(ng);
        for (int64_t i = 0; i < ng; ++i) vn[i] = (float)(sd > 0.0 ? (V[i] - mean) / sd : 0.0);
        c->vnorm.upload(vn.data(), ng);
        // weights: seeded normal draws, scaled so that the initial correction is small
        std::mt19937_64 rng(seed);
        std::normal_distribution<double> nd(0.0, 1.0);
        c->params.assign(kNumParams, 0.f);
        for (int k = kOffW1; k < kOffW2; ++k) c->params[k] = (float)(0.3 / std::sqrt(54.0) * nd(rng));
        for (int k = kOffW2; k < kOffW3; ++k) c->params[k] = (float)(0.3 / std::sqrt(216.0) * nd(rng));
        for (int k = kOffW3; k < kOffA1; ++k) c->params[k] = (float)(0.3 / std::sqrt(216.0) * nd(rng));
        c->p.upload(c->params.data(), kNumParams);
        c->grad.alloc(kNumParams);
        c->m.alloc(kNumParams);
        c->v2.alloc(kNumParams);
        QG_HIP(hipMemset(c->m.p, 0, kNumParams * 4));
        QG_HIP(hipMemset(c->v2.p, 0, kNumParams * 4));
        if (max_pairs > 0) {
            c->U.alloc((size_t)max_pairs * ng);
            c->Tt.alloc((size_t)max_pairs * ng);
        }
        *out = c.release();
    });
}

int qg_cnn_destroy(qg_cnn* c) {
    return guarded([&] {
        if (c && c->h && c->h->cnn == c) c->h->cnn = nullptr;
        delete c;
    });
}

int qg_cnn_get_params(qg_cnn* c, float* host, int n) {
    return guarded([&] {
        QG_REQUIRE(c && host && n == kNumParams, "bad arguments");
        QG_HIP(hipMemcpy(host, c->p.p, kNumParams * 4, hipMemcpyDeviceToHost));
    });
}

int qg_cnn_set_params(qg_cnn* c, const float* host, int n) {
    return guarded([&] {
        QG_REQUIRE(c && host && n == kNumParams, "bad arguments");
        std::copy(host, host + n, c->params.begin());
        QG_HIP(hipMemcpy(c->p.p, host, kNumParams * 4, hipMemcpyHostToDevice));
        QG_HIP(hipMemset(c->m.p, 0, kNumParams * 4));
        QG_HIP(hipMemset(c->v2.p, 0, kNumParams * 4));
        c->adam_t = 0;
        c->dirty = true;
    });
}

int qg_cnn_network(qg_cnn* c, int adjoint, int bf16, int nbands, const float* in, float* out, void* stream) {
    return guarded([&] {
        QG_REQUIRE(c && in && out && nbands >= 0, "bad arguments");
        if (nbands == 0) return;
        hipStream_t s = as_stream(stream);
        refresh_derived(c, s);
        const Weights W = weights(c, bf16 != 0);
        const int64_t ng = c->h->grid.ng();
        DevBuf<float> s8;
        s8.alloc((size_t)2 * nbands * kC * ng);
        float* a = s8.p;
        float* b = s8.p + (int64_t)nbands * kC * ng;
        if (bf16 == 2) {
            if (adjoint) net_adjoint_mfma(c, W, nbands, in, out, a, b, s);
            else net_forward_mfma(c, W, nbands, in, out, a, b, s);
        } else if (adjoint) {
            net_adjoint(c, W, nbands, in, out, a, b, nullptr, nullptr, s);
        } else {
            net_forward(c, W, nbands, in, out, nullptr, a, b, s);
        }
        QG_HIP(hipStreamSynchronize(s));
    });
}

int qg_cnn_loss_grad(qg_cnn* c, int nbands, const float* u, const float* t, double znorm, double* loss,
                     float* grad_host) {
    return guarded([&] {
        QG_REQUIRE(c && u && t && nbands >= 1 && znorm > 0.0, "bad arguments");
        hipStream_t s = nullptr;
        const double l = loss_grad(c, nbands, u, t, znorm, s);
        if (loss) *loss = l;
        if (grad_host) QG_HIP(hipMemcpy(grad_host, c->grad.p, kNumParams * 4, hipMemcpyDeviceToHost));
    });
}

int qg_cnn_collect(qg_cnn* c, int nbands, const double* X, const double* eig_host, double sigma, double tol,
                   int maxit, int* cg_iters) {
    return guarded([&] {
        QG_REQUIRE(c && X && eig_host && nbands >= 1 && tol > 0.0 && maxit > 0, "bad arguments");
        const int it = collect_pairs(c, nbands, X, eig_host, sigma, tol, maxit, nullptr);
        if (cg_iters) *cg_iters = it;
    });
}

int qg_cnn_num_pairs(const qg_cnn* c) { return c ? c->npairs : -1; }

int qg_cnn_train(qg_cnn* c, int steps, int batch, double lr, uint64_t seed, double* loss_first, double* loss_last) {
    return guarded([&] {
        QG_REQUIRE(c && steps >= 1 && batch >= 1 && lr > 0.0, "bad arguments");
        train(c, steps, batch, lr, seed, loss_first, loss_last, nullptr);
    });
}

int qg_cnn_apply(qg_cnn* c, int dtype, int nbands, void* r, int64_t ld, void* stream) {
    return guarded([&] {
        QG_REQUIRE(c && r && nbands >= 0, "bad arguments");
        QG_REQUIRE(dtype == QG_F64 || dtype == QG_F32, "bad dtype");
        if (nbands == 0) return;
        cnn_precond_apply(c, dtype, nbands, r, ld, as_stream(stream));
    });
}

int qg_hamiltonian_set_cnn(qg_ham* h, qg_cnn* c) {
    return guarded([&] {
        QG_REQUIRE(h, "null Hamiltonian");
        QG_REQUIRE(!c || c->h == h, "the network was created for another Hamiltonian");
        h->cnn = c;
    });
}

}  // extern "C"
