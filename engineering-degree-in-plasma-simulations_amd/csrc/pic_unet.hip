// pic_unet.hip -- learned warm start and preconditioner for the Poisson solve
// (no counterpart in the reference).
//
// Network: a 3-level 3-D convolutional U-Net on the node grid.  Activations
// are channels-last BF16 ([voxel][C]); weights are BF16 for the MFMA layers
// with FP32 master copies; every accumulation is FP32.
//
//   L1 (n^3)   a1a = relu(conv3(x, 2->C1))            direct kernel (FP32 input)
//              s1  = relu(conv3(a1a, C1->C1))         MFMA
//   L2 (n/2)   a2a = relu(conv3(pool(s1), C1->C2))    MFMA
//              s2  = relu(conv3(a2a, C2->C2))         MFMA
//   L3 (n/4)   a3a = relu(conv3(pool(s2), C2->C3))    MFMA
//              a3b = relu(conv3(a3a, C3->C3))         MFMA
//   L2         a2  = up(conv1(a3b, C3->C2)) + s2      MFMA (1x1) + nearest upsample
//              c2c = relu(conv3(a2, C2->C2))          MFMA
//   L1         a1  = up(conv1(c2c, C2->C1)) + s1
//              c1c = relu(conv3(a1, C1->C1))          MFMA
//              y   = conv3(c1c, C1->1)                direct kernel (FP32 output)
//
// Input channels: the right-hand side in grid units (normalised) and the
// Dirichlet mask.  Target: the potential of the unit-spacing Poisson problem
// (laplacian(phi) = -rho, phi = 0 on Dirichlet nodes).
//
// MFMA convolution (implicit GEMM): a workgroup owns 16 x 4 x 4 output
// voxels (16 M-blocks of 16 voxels along x); the input tile plus a one-voxel
// halo is staged in LDS 16 channels at a time; each of the 4 waves issues
// v_mfma_f32_16x16x16_bf16 for its 4 M-blocks and up to 4 N-blocks of 16
// output channels, K = 16 input channels of one tap.  Backward-data is the
// same kernel with flipped taps and transposed channels, ReLU masks fused in
// the epilogue.  Weight gradients are register-blocked FP32 reductions over
// voxel tiles staged in LDS.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "pic_internal.h"

using namespace pic;

namespace {

typedef uint16_t bf16;
typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __host__ inline float bf2f(bf16 h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
__device__ __host__ inline bf16 f2bf(float f) {  // round to nearest even
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}

struct Dims {
    int nx, ny, nz;
    long long v;
};
Dims half_dims(const Dims& d) {
    Dims h{d.nx / 2, d.ny / 2, d.nz / 2, 0};
    h.v = (long long)h.nx * h.ny * h.nz;
    return h;
}

__device__ inline void vox_xyz(const Dims& d, long long v, int& x, int& y, int& z) {
    x = (int)(v % d.nx);
    long long r = v / d.nx;
    y = (int)(r % d.ny);
    z = (int)(r / d.ny);
}

// ---------------------------------------------------------------- MFMA conv
constexpr int TX = 16, TY = 4, TZ = 4;
constexpr int NBB = 4;  // N-blocks (16 output channels each) per workgroup

struct ConvArgs {
    const bf16* in;
    int cin;
    const bf16* w;  // [cout][taps * cin], k = tap * cin + ci
    const float* bias;
    bf16* out;
    int cout;
    const bf16* mask;  // optional: keep outputs where mask[v][n] > 0 (mask has cout channels)
    int relu, accumulate;
    Dims d;
};

template <int TAPS>
__global__ void __launch_bounds__(256) k_conv_mfma(ConvArgs a) {
    constexpr int H = TAPS == 27 ? 1 : 0;
    constexpr int SX = TX + 2 * H, SY = TY + 2 * H, SZ = TZ + 2 * H;
    __shared__ uint4 tile[SX * SY * SZ * 2];  // 16 BF16 channels = 32 bytes per voxel
    const Dims d = a.d;
    const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
    const int b = blockIdx.x;
    const int x0 = (b % ntx) * TX, y0 = ((b / ntx) % nty) * TY, z0 = (b / (ntx * nty)) * TZ;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, m = lane & 15;
    const int nb0 = blockIdx.y * NBB;
    const int nbc = min(NBB, a.cout / 16 - nb0);
    f4v acc[4][NBB];
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int nb = 0; nb < NBB; nb++) acc[mb][nb] = f4v{0.f, 0.f, 0.f, 0.f};
    for (int c0 = 0; c0 < a.cin; c0 += 16) {
        __syncthreads();
        for (int s = tid; s < SX * SY * SZ * 2; s += 256) {
            int vox = s >> 1, hf = s & 1;
            int sx = vox % SX, sy = (vox / SX) % SY, sz = vox / (SX * SY);
            int gx = x0 + sx - H, gy = y0 + sy - H, gz = z0 + sz - H;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
                val = *reinterpret_cast<const uint4*>(
                    a.in + (((long long)gz * d.ny + gy) * d.nx + gx) * a.cin + c0 + hf * 8);
            tile[s] = val;
        }
        __syncthreads();
        for (int t = 0; t < TAPS; t++) {
            const int ox = TAPS == 27 ? t % 3 : 0, oy = TAPS == 27 ? (t / 3) % 3 : 0,
                      oz = TAPS == 27 ? t / 9 : 0;
            s4v bfr[NBB];
#pragma unroll
            for (int nb = 0; nb < NBB; nb++) {
                if (nb < nbc) {
                    int n = (nb0 + nb) * 16 + m;
                    bfr[nb] = *reinterpret_cast<const s4v*>(a.w + (long long)n * TAPS * a.cin +
                                                            t * a.cin + c0 + 4 * g);
                }
            }
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
                int row = wave * 4 + mb;
                int ty = row % TY, tz = row / TY;
                int sv = ((tz + oz) * SY + (ty + oy)) * SX + (m + ox);
                s4v afr = *reinterpret_cast<const s4v*>(reinterpret_cast<const char*>(tile) +
                                                        sv * 32 + 8 * g);
#pragma unroll
                for (int nb = 0; nb < NBB; nb++)
                    if (nb < nbc)
                        acc[mb][nb] =
                            __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(afr, bfr[nb], acc[mb][nb], 0, 0, 0);
            }
        }
    }
    // D fragment: lane holds rows 4g + j (x offset) of column m (channel)
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        int row = wave * 4 + mb;
        int gy = y0 + row % TY, gz = z0 + row / TY;
        if (gy >= d.ny || gz >= d.nz) continue;
#pragma unroll
        for (int nb = 0; nb < NBB; nb++) {
            if (nb >= nbc) continue;
            int n = (nb0 + nb) * 16 + m;
            float bv = a.bias ? a.bias[n] : 0.f;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int gx = x0 + 4 * g + j;
                if (gx >= d.nx) continue;
                long long o = (((long long)gz * d.ny + gy) * d.nx + gx) * a.cout + n;
                float v = acc[mb][nb][j] + bv;
                if (a.relu) v = fmaxf(v, 0.f);
                if (a.mask && !(bf2f(a.mask[o]) > 0.f)) v = 0.f;
                if (a.accumulate) v += bf2f(a.out[o]);
                a.out[o] = f2bf(v);
            }
        }
    }
}

// ------------------------------------------------------- direct conv layers
// input layer: FP32 [V][2] -> BF16 [V][C] with bias and ReLU
__global__ void __launch_bounds__(256)
    k_conv_in(Dims d, const float* __restrict__ x, const float* __restrict__ w,
              const float* __restrict__ bias, bf16* __restrict__ out, int C) {
    __shared__ float ws[32 * 54];
    for (int t = threadIdx.x; t < C * 54; t += 256) ws[t] = w[t];
    __syncthreads();
    long long v = (long long)blockIdx.x * 256 + threadIdx.x;
    if (v >= d.v) return;
    int xi, yi, zi;
    vox_xyz(d, v, xi, yi, zi);
    float acc[32];
    for (int c = 0; c < C; c++) acc[c] = bias[c];
    for (int t = 0; t < 27; t++) {
        int gx = xi + t % 3 - 1, gy = yi + (t / 3) % 3 - 1, gz = zi + t / 9 - 1;
        if (gx < 0 || gx >= d.nx || gy < 0 || gy >= d.ny || gz < 0 || gz >= d.nz) continue;
        float2 in = *reinterpret_cast<const float2*>(x + 2 * (((long long)gz * d.ny + gy) * d.nx + gx));
        for (int c = 0; c < C; c++) acc[c] += in.x * ws[c * 54 + 2 * t] + in.y * ws[c * 54 + 2 * t + 1];
    }
    for (int c = 0; c < C; c++) out[v * C + c] = f2bf(fmaxf(acc[c], 0.f));
}

// output layer: BF16 [V][C] -> FP32 [V], bias, no activation; w is [27][C]
__global__ void __launch_bounds__(256)
    k_conv_out(Dims d, const bf16* __restrict__ x, int C, const float* __restrict__ w,
               const float* __restrict__ bias, float* __restrict__ y) {
    __shared__ float ws[27 * 32];
    for (int t = threadIdx.x; t < 27 * C; t += 256) ws[t] = w[t];
    __syncthreads();
    long long v = (long long)blockIdx.x * 256 + threadIdx.x;
    if (v >= d.v) return;
    int xi, yi, zi;
    vox_xyz(d, v, xi, yi, zi);
    float acc = bias[0];
    for (int t = 0; t < 27; t++) {
        int gx = xi + t % 3 - 1, gy = yi + (t / 3) % 3 - 1, gz = zi + t / 9 - 1;
        if (gx < 0 || gx >= d.nx || gy < 0 || gy >= d.ny || gz < 0 || gz >= d.nz) continue;
        const bf16* row = x + (((long long)gz * d.ny + gy) * d.nx + gx) * C;
        for (int c = 0; c < C; c++) acc += bf2f(row[c]) * ws[t * C + c];
    }
    y[v] = acc;
}

// d(input of the output layer) = sum_t dy[v - off t] w[t][c], times the ReLU
// mask of that input
__global__ void __launch_bounds__(256)
    k_conv_out_bwd(Dims d, const float* __restrict__ dy, int C, const float* __restrict__ w,
                   const bf16* __restrict__ act, bf16* __restrict__ dx) {
    __shared__ float ws[27 * 32];
    for (int t = threadIdx.x; t < 27 * C; t += 256) ws[t] = w[t];
    __syncthreads();
    long long v = (long long)blockIdx.x * 256 + threadIdx.x;
    if (v >= d.v) return;
    int xi, yi, zi;
    vox_xyz(d, v, xi, yi, zi);
    float acc[32];
    for (int c = 0; c < C; c++) acc[c] = 0.f;
    for (int t = 0; t < 27; t++) {
        int gx = xi - (t % 3 - 1), gy = yi - ((t / 3) % 3 - 1), gz = zi - (t / 9 - 1);
        if (gx < 0 || gx >= d.nx || gy < 0 || gy >= d.ny || gz < 0 || gz >= d.nz) continue;
        float g = dy[((long long)gz * d.ny + gy) * d.nx + gx];
        for (int c = 0; c < C; c++) acc[c] += g * ws[t * C + c];
    }
    for (int c = 0; c < C; c++) {
        bool on = bf2f(act[v * C + c]) > 0.f;
        dx[v * C + c] = f2bf(on ? acc[c] : 0.f);
    }
}

// ------------------------------------------------------ pooling / upsampling
__global__ void __launch_bounds__(256)
    k_pool(Dims dc, const bf16* __restrict__ fine, bf16* __restrict__ coarse, int C) {
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= dc.v * C) return;
    long long v = e / C;
    int c = (int)(e - v * C);
    int xi, yi, zi;
    vox_xyz(dc, v, xi, yi, zi);
    int fx = 2 * dc.nx, fy = 2 * dc.ny;
    float s = 0.f;
    for (int q = 0; q < 8; q++) {
        long long fv = ((long long)(2 * zi + (q >> 2)) * fy + (2 * yi + ((q >> 1) & 1))) * fx +
                       (2 * xi + (q & 1));
        s += bf2f(fine[fv * C + c]);
    }
    coarse[e] = f2bf(0.125f * s);
}

// d_fine = (d_fine + d_coarse[parent] / 8) * (act > 0)
__global__ void __launch_bounds__(256)
    k_pool_bwd(Dims df, const bf16* __restrict__ dcoarse, bf16* __restrict__ dfine,
               const bf16* __restrict__ act, int C) {
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= df.v * C) return;
    long long v = e / C;
    int c = (int)(e - v * C);
    int xi, yi, zi;
    vox_xyz(df, v, xi, yi, zi);
    long long pv = ((long long)(zi >> 1) * (df.ny >> 1) + (yi >> 1)) * (df.nx >> 1) + (xi >> 1);
    float g = bf2f(dfine[e]) + 0.125f * bf2f(dcoarse[pv * C + c]);
    dfine[e] = f2bf(bf2f(act[e]) > 0.f ? g : 0.f);
}

// fine = up(coarse) + skip
__global__ void __launch_bounds__(256)
    k_up_add(Dims df, const bf16* __restrict__ coarse, const bf16* __restrict__ skip,
             bf16* __restrict__ fine, int C) {
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= df.v * C) return;
    long long v = e / C;
    int c = (int)(e - v * C);
    int xi, yi, zi;
    vox_xyz(df, v, xi, yi, zi);
    long long pv = ((long long)(zi >> 1) * (df.ny >> 1) + (yi >> 1)) * (df.nx >> 1) + (xi >> 1);
    fine[e] = f2bf(bf2f(coarse[pv * C + c]) + bf2f(skip[e]));
}

// d_coarse = sum over the 8 children of d_fine
__global__ void __launch_bounds__(256)
    k_up_bwd(Dims dc, const bf16* __restrict__ dfine, bf16* __restrict__ dcoarse, int C) {
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= dc.v * C) return;
    long long v = e / C;
    int c = (int)(e - v * C);
    int xi, yi, zi;
    vox_xyz(dc, v, xi, yi, zi);
    int fx = 2 * dc.nx, fy = 2 * dc.ny;
    float s = 0.f;
    for (int q = 0; q < 8; q++) {
        long long fv = ((long long)(2 * zi + (q >> 2)) * fy + (2 * yi + ((q >> 1) & 1))) * fx +
                       (2 * xi + (q & 1));
        s += bf2f(dfine[fv * C + c]);
    }
    dcoarse[e] = f2bf(s);
}

// ------------------------------------------------------- weight gradients
// MFMA layers: dW[co][t*cin + ci] += sum_v dY[v][co] X[v + off t][ci].
// A workgroup stages 8x4x4 voxel tiles of dY and (tile + halo) of X in LDS
// and loops over tiles_per_block tiles; each thread owns up to 4 register
// blocks of 4 (co) x 4 (ci) for one tap.
constexpr int WX = 8, WY = 4, WZ = 4;
template <int TAPS>
__global__ void __launch_bounds__(256)
    k_wgrad(Dims d, const bf16* __restrict__ X, int cin, const bf16* __restrict__ dY, int cout,
            float* __restrict__ dW, int tiles_per_block) {
    constexpr int H = TAPS == 27 ? 1 : 0;
    constexpr int SX = WX + 2 * H, SY = WY + 2 * H, SZ = WZ + 2 * H;
    extern __shared__ bf16 lds[];
    bf16* xs = lds;                       // [SX*SY*SZ][cin]
    bf16* ys = lds + SX * SY * SZ * cin;  // [WX*WY*WZ][cout]
    const int ntx = (d.nx + WX - 1) / WX, nty = (d.ny + WY - 1) / WY, ntz = (d.nz + WZ - 1) / WZ;
    const int ntiles = ntx * nty * ntz;
    const int ncob = cout / 4, ncib = cin / 4;
    const int combos = TAPS * ncob * ncib;
    float acc[4][16];
    int cidx[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        cidx[r] = blockIdx.y * 1024 + threadIdx.x + 256 * r;
#pragma unroll
        for (int q = 0; q < 16; q++) acc[r][q] = 0.f;
    }
    for (int tt = 0; tt < tiles_per_block; tt++) {
        int tile = blockIdx.x * tiles_per_block + tt;
        if (tile >= ntiles) break;
        int x0 = (tile % ntx) * WX, y0 = ((tile / ntx) % nty) * WY, z0 = (tile / (ntx * nty)) * WZ;
        __syncthreads();
        for (int s = threadIdx.x; s < SX * SY * SZ * cin; s += 256) {
            int vox = s / cin, c = s - vox * cin;
            int sx = vox % SX, sy = (vox / SX) % SY, sz = vox / (SX * SY);
            int gx = x0 + sx - H, gy = y0 + sy - H, gz = z0 + sz - H;
            bf16 val = 0;
            if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
                val = X[(((long long)gz * d.ny + gy) * d.nx + gx) * cin + c];
            xs[s] = val;
        }
        for (int s = threadIdx.x; s < WX * WY * WZ * cout; s += 256) {
            int vox = s / cout, c = s - vox * cout;
            int sx = vox % WX, sy = (vox / WX) % WY, sz = vox / (WX * WY);
            int gx = x0 + sx, gy = y0 + sy, gz = z0 + sz;
            bf16 val = 0;
            if (gx < d.nx && gy < d.ny && gz < d.nz)
                val = dY[(((long long)gz * d.ny + gy) * d.nx + gx) * cout + c];
            ys[s] = val;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int ci_ = cidx[r];
            if (ci_ >= combos) continue;
            int t = ci_ / (ncob * ncib);
            int rem = ci_ - t * ncob * ncib;
            int cob = rem / ncib, cib = rem - cob * ncib;
            int ox = TAPS == 27 ? t % 3 : 0, oy = TAPS == 27 ? (t / 3) % 3 : 0,
                oz = TAPS == 27 ? t / 9 : 0;
            for (int v = 0; v < WX * WY * WZ; v++) {
                int vx = v % WX, vy = (v / WX) % WY, vz = v / (WX * WY);
                const bf16* xr = xs + (((vz + oz) * SY + (vy + oy)) * SX + (vx + ox)) * cin + 4 * cib;
                const bf16* yr = ys + v * cout + 4 * cob;
                float xv[4], yv[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    xv[q] = bf2f(xr[q]);
                    yv[q] = bf2f(yr[q]);
                }
#pragma unroll
                for (int p = 0; p < 4; p++)
#pragma unroll
                    for (int q = 0; q < 4; q++) acc[r][p * 4 + q] += yv[p] * xv[q];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        int ci_ = cidx[r];
        if (ci_ >= combos) continue;
        int t = ci_ / (ncob * ncib);
        int rem = ci_ - t * ncob * ncib;
        int cob = rem / ncib, cib = rem - cob * ncib;
#pragma unroll
        for (int p = 0; p < 4; p++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                atomicAdd(&dW[(long long)(4 * cob + p) * TAPS * cin + t * cin + 4 * cib + q],
                          acc[r][p * 4 + q]);
    }
}

// direct layers' weight gradients: one thread per weight, a chunk of voxels
// per workgroup.  in layer: dW[c][2t + i] = sum dY[v][c] x[v + off t][i];
// out layer: dW[t][c] = sum dy[v] X[v + off t][c].
constexpr int DCHUNK = 4096;
__global__ void __launch_bounds__(256)
    k_wgrad_in(Dims d, const float* __restrict__ x, const bf16* __restrict__ dY, int C,
               float* __restrict__ dW, float* __restrict__ db) {
    long long v0 = (long long)blockIdx.x * DCHUNK;
    long long v1 = min(v0 + DCHUNK, d.v);
    for (int wi = threadIdx.x; wi < C * 55; wi += 256) {
        int c = wi / 55, k = wi - c * 55;
        float acc = 0.f;
        if (k == 54) {
            for (long long v = v0; v < v1; v++) acc += bf2f(dY[v * C + c]);
            atomicAdd(&db[c], acc);
            continue;
        }
        int t = k >> 1, ch = k & 1;
        int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = t / 9 - 1;
        for (long long v = v0; v < v1; v++) {
            int xi, yi, zi;
            vox_xyz(d, v, xi, yi, zi);
            int gx = xi + ox, gy = yi + oy, gz = zi + oz;
            if (gx < 0 || gx >= d.nx || gy < 0 || gy >= d.ny || gz < 0 || gz >= d.nz) continue;
            acc += bf2f(dY[v * C + c]) * x[2 * (((long long)gz * d.ny + gy) * d.nx + gx) + ch];
        }
        atomicAdd(&dW[c * 54 + k], acc);
    }
}

__global__ void __launch_bounds__(256)
    k_wgrad_out(Dims d, const bf16* __restrict__ X, int C, const float* __restrict__ dy,
                float* __restrict__ dW, float* __restrict__ db) {
    long long v0 = (long long)blockIdx.x * DCHUNK;
    long long v1 = min(v0 + DCHUNK, d.v);
    for (int wi = threadIdx.x; wi <= 27 * C; wi += 256) {
        float acc = 0.f;
        if (wi == 27 * C) {
            for (long long v = v0; v < v1; v++) acc += dy[v];
            atomicAdd(db, acc);
            continue;
        }
        int t = wi / C, c = wi - t * C;
        int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = t / 9 - 1;
        for (long long v = v0; v < v1; v++) {
            int xi, yi, zi;
            vox_xyz(d, v, xi, yi, zi);
            int gx = xi + ox, gy = yi + oy, gz = zi + oz;
            if (gx < 0 || gx >= d.nx || gy < 0 || gy >= d.ny || gz < 0 || gz >= d.nz) continue;
            acc += dy[v] * bf2f(X[(((long long)gz * d.ny + gy) * d.nx + gx) * C + c]);
        }
        atomicAdd(&dW[wi], acc);
    }
}

// bias gradient of an MFMA layer: db[c] += sum_v dY[v][c]
__global__ void __launch_bounds__(256)
    k_bias_grad(long long V, const bf16* __restrict__ dY, int C, float* __restrict__ db) {
    int c = threadIdx.x % C;
    int lanes = 256 / C * C;
    if (threadIdx.x >= lanes) return;
    float acc = 0.f;
    long long step = (long long)gridDim.x * (lanes / C);
    for (long long v = (long long)blockIdx.x * (lanes / C) + threadIdx.x / C; v < V; v += step)
        acc += bf2f(dY[v * C + c]);
    atomicAdd(&db[c], acc);
}

// ---------------------------------------------------------- misc kernels
// transposed, tap-flipped BF16 weights for backward-data:
// wt[ci][t' * cout + co] = w[co][(T-1-t') * cin + ci]
__global__ void k_transpose_w(const float* __restrict__ w, int cout, int cin, int taps,
                              bf16* __restrict__ wb, bf16* __restrict__ wt) {
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    long long n = (long long)cout * taps * cin;
    if (e >= n) return;
    int co = (int)(e / (taps * cin));
    int rem = (int)(e - (long long)co * taps * cin);
    int t = rem / cin, ci = rem - t * cin;
    bf16 v = f2bf(w[e]);
    wb[e] = v;
    wt[(long long)ci * taps * cout + (long long)(taps - 1 - t) * cout + co] = v;
}

// loss over non-Dirichlet voxels: dy = (y - t) / Nvalid, out[0] += 0.5 sum (y-t)^2 / Nvalid
__global__ void __launch_bounds__(256)
    k_loss(long long V, const float* __restrict__ y, const float* __restrict__ t,
           const float* __restrict__ x2, float inv_n, float* __restrict__ dy,
           double* __restrict__ out) {
    long long v = (long long)blockIdx.x * 256 + threadIdx.x;
    double l = 0;
    if (v < V) {
        float valid = x2[2 * v + 1] > 0.5f ? 0.f : 1.f;
        float r = (y[v] - t[v]) * valid;
        dy[v] = r * inv_n;
        l = 0.5 * (double)r * r * inv_n;
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, l);
}

__global__ void __launch_bounds__(256)
    k_adam(long long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
           float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
           float gscale) {
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    float gr = g[e] * gscale;
    float mm = b1 * m[e] + (1 - b1) * gr;
    float vv = b2 * v[e] + (1 - b2) * gr * gr;
    m[e] = mm;
    v[e] = vv;
    p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

}  // namespace

// =========================================================================
// host side: network state
// =========================================================================
struct UNetLayer {
    int cin, cout, taps;  // taps 27 or 1; direct layers: in (cin 2) / out (cout 1)
    size_t woff, boff;    // offsets into the FP32 parameter vector
    bf16* wb = nullptr;   // BF16 weights [cout][taps*cin]
    bf16* wt = nullptr;   // BF16 transposed, tap-flipped
};

struct UNet {
    int C1, C2, C3;
    Dims d1, d2, d3;
    std::vector<UNetLayer> L;  // order: in, c1b, c2a, c2b, c3a, c3b, u2, c2c, u1, c1c, out
    size_t nparam = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    // activations (BF16 channels-last) and their gradients
    float* x = nullptr;   // [V1][2]
    float* y = nullptr;   // [V1]
    float* dy = nullptr;  // [V1]
    bf16 *a1a, *s1, *p1, *a2a, *s2, *p2, *a3a, *a3b, *u2, *a2, *c2c, *u1, *a1, *c1c;
    bf16 *g1a, *gs1, *gp1, *g2a, *gs2, *gp2, *g3a, *g3b, *gu2, *g2, *gc2c, *gu1, *g1, *gc1c;
    long long adam_t = 0;
    std::vector<void*> allocs;
    bool trained = false;
};

enum { L_IN = 0, L_C1B, L_C2A, L_C2B, L_C3A, L_C3B, L_U2, L_C2C, L_U1, L_C1C, L_OUT, L_N };

void pic_unet_free(pic_ctx* ctx) {
    UNet* u = (UNet*)ctx->unet;
    if (!u) return;
    for (void* p : u->allocs) (void)hipFree(p);
    delete u;
    ctx->unet = nullptr;
}

static int ualloc(pic_ctx* ctx, UNet* u, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return pic_hip_check(ctx, e, "U-Net allocation");
    u->allocs.push_back(*p);
    return PIC_OK;
}

// deterministic He-style initialisation (host, then uploaded)
static void init_params(UNet* u, uint32_t seed, std::vector<float>& h) {
    h.assign(u->nparam, 0.f);
    uint32_t s = seed * 2654435761u + 12345u;
    auto rnd = [&]() {  // xorshift32 -> uniform(-1, 1)
        s ^= s << 13;
        s ^= s >> 17;
        s ^= s << 5;
        return (float)((s >> 8) * (1.0 / 8388608.0) - 1.0);
    };
    for (auto& l : u->L) {
        size_t nw = (size_t)l.cout * l.taps * l.cin;
        float scale = std::sqrt(6.0f / (float)(l.taps * l.cin));
        if (&l == &u->L[L_OUT]) scale *= 0.1f;
        for (size_t q = 0; q < nw; q++) h[l.woff + q] = scale * rnd();
    }
}

static int unet_build(pic_ctx* ctx, int C1, int C2, int C3, uint32_t seed) {
    pic_unet_free(ctx);
    const GridC& g = ctx->g;
    if (g.ni % 4 || g.nj % 4 || g.nk % 4)
        return pic_fail(ctx, PIC_ERR_ARG, "U-Net needs node counts divisible by 4");
    if (C1 % 16 || C2 % 16 || C3 % 16 || C1 > 32 || C2 > 64 || C3 > 64)
        return pic_fail(ctx, PIC_ERR_ARG, "U-Net channels must be multiples of 16 (C1<=32, C2,C3<=64)");
    UNet* u = new UNet();
    ctx->unet = u;
    u->C1 = C1;
    u->C2 = C2;
    u->C3 = C3;
    u->d1 = Dims{g.ni, g.nj, g.nk, g.nn};
    u->d2 = half_dims(u->d1);
    u->d3 = half_dims(u->d2);
    int spec[L_N][3] = {{2, C1, 27},  {C1, C1, 27}, {C1, C2, 27}, {C2, C2, 27},
                        {C2, C3, 27}, {C3, C3, 27}, {C3, C2, 1},  {C2, C2, 27},
                        {C2, C1, 1},  {C1, C1, 27}, {C1, 1, 27}};
    size_t off = 0;
    for (int l = 0; l < L_N; l++) {
        UNetLayer ly;
        ly.cin = spec[l][0];
        ly.cout = spec[l][1];
        ly.taps = spec[l][2];
        ly.woff = off;
        off += (size_t)ly.cin * ly.cout * ly.taps;
        ly.boff = off;
        off += ly.cout;
        u->L.push_back(ly);
    }
    u->nparam = off;
    int rc = 0;
    size_t pb = off * sizeof(float);
    rc |= ualloc(ctx, u, (void**)&u->p, pb);
    rc |= ualloc(ctx, u, (void**)&u->g, pb);
    rc |= ualloc(ctx, u, (void**)&u->m, pb);
    rc |= ualloc(ctx, u, (void**)&u->v, pb);
    for (int l = L_C1B; l < L_OUT; l++) {
        size_t nw = (size_t)u->L[l].cin * u->L[l].cout * u->L[l].taps * sizeof(bf16);
        rc |= ualloc(ctx, u, (void**)&u->L[l].wb, nw);
        rc |= ualloc(ctx, u, (void**)&u->L[l].wt, nw);
    }
    long long V1 = u->d1.v, V2 = u->d2.v, V3 = u->d3.v;
    rc |= ualloc(ctx, u, (void**)&u->x, V1 * 2 * sizeof(float));
    rc |= ualloc(ctx, u, (void**)&u->y, V1 * sizeof(float));
    rc |= ualloc(ctx, u, (void**)&u->dy, V1 * sizeof(float));
    struct A {
        bf16** p;
        long long n;
    } acts[] = {{&u->a1a, V1 * C1}, {&u->s1, V1 * C1},  {&u->p1, V2 * C1},  {&u->a2a, V2 * C2},
                {&u->s2, V2 * C2},  {&u->p2, V3 * C2},  {&u->a3a, V3 * C3}, {&u->a3b, V3 * C3},
                {&u->u2, V3 * C2},  {&u->a2, V2 * C2},  {&u->c2c, V2 * C2}, {&u->u1, V2 * C1},
                {&u->a1, V1 * C1},  {&u->c1c, V1 * C1}, {&u->g1a, V1 * C1}, {&u->gs1, V1 * C1},
                {&u->gp1, V2 * C1}, {&u->g2a, V2 * C2}, {&u->gs2, V2 * C2}, {&u->gp2, V3 * C2},
                {&u->g3a, V3 * C3}, {&u->g3b, V3 * C3}, {&u->gu2, V3 * C2}, {&u->g2, V2 * C2},
                {&u->gc2c, V2 * C2}, {&u->gu1, V2 * C1}, {&u->g1, V1 * C1}, {&u->gc1c, V1 * C1}};
    for (auto& a : acts) rc |= ualloc(ctx, u, (void**)a.p, a.n * sizeof(bf16));
    if (rc) {
        pic_unet_free(ctx);
        return PIC_ERR_HIP;
    }
    std::vector<float> h;
    init_params(u, seed, h);
    PIC_HIP(ctx, hipMemcpyAsync(u->p, h.data(), pb, hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP(ctx, hipMemsetAsync(u->m, 0, pb, ctx->stream));
    PIC_HIP(ctx, hipMemsetAsync(u->v, 0, pb, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

// refresh the BF16 weight copies from the FP32 master parameters
static int unet_sync_weights(pic_ctx* ctx, UNet* u) {
    for (int l = L_C1B; l < L_OUT; l++) {
        UNetLayer& ly = u->L[l];
        long long n = (long long)ly.cout * ly.taps * ly.cin;
        k_transpose_w<<<grid_for(n, 256), 256, 0, ctx->stream>>>(u->p + ly.woff, ly.cout, ly.cin,
                                                                ly.taps, ly.wb, ly.wt);
    }
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

static void conv(pic_ctx* ctx, const Dims& d, const bf16* in, int cin, const bf16* w,
                 const float* bias, bf16* out, int cout, int taps, int relu, const bf16* mask,
                 int accumulate) {
    ConvArgs a{in, cin, w, bias, out, cout, mask, relu, accumulate, d};
    int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY, ntz = (d.nz + TZ - 1) / TZ;
    dim3 grid(ntx * nty * ntz, (cout / 16 + NBB - 1) / NBB);
    if (taps == 27)
        k_conv_mfma<27><<<grid, 256, 0, ctx->stream>>>(a);
    else
        k_conv_mfma<1><<<grid, 256, 0, ctx->stream>>>(a);
}

static void wgrad(pic_ctx* ctx, const Dims& d, const bf16* X, int cin, const bf16* dY, int cout,
                  int taps, float* dW, float* db) {
    int ntx = (d.nx + WX - 1) / WX, nty = (d.ny + WY - 1) / WY, ntz = (d.nz + WZ - 1) / WZ;
    int ntiles = ntx * nty * ntz;
    int tpb = std::max(1, ntiles / 512);
    int combos = taps * (cout / 4) * (cin / 4);
    dim3 grid((ntiles + tpb - 1) / tpb, (combos + 1023) / 1024);
    int H = taps == 27 ? 1 : 0;
    size_t lds = ((size_t)(WX + 2 * H) * (WY + 2 * H) * (WZ + 2 * H) * cin +
                  (size_t)WX * WY * WZ * cout) * sizeof(bf16);
    if (taps == 27)
        k_wgrad<27><<<grid, 256, lds, ctx->stream>>>(d, X, cin, dY, cout, dW, tpb);
    else
        k_wgrad<1><<<grid, 256, lds, ctx->stream>>>(d, X, cin, dY, cout, dW, tpb);
    unsigned nb = (unsigned)std::min<long long>(1024, (d.v + 255) / 256);
    k_bias_grad<<<nb, 256, 0, ctx->stream>>>(d.v, dY, cout, db);
}

static unsigned ew(long long n) { return grid_for((size_t)n, 256); }

// forward pass: u->x (FP32 [V1][2]) -> u->y (FP32 [V1])
static int unet_forward(pic_ctx* ctx, UNet* u) {
    const Dims &d1 = u->d1, &d2 = u->d2, &d3 = u->d3;
    const int C1 = u->C1, C2 = u->C2, C3 = u->C3;
    auto& L = u->L;
    float* P = u->p;
    PicProfScope ps(ctx, PIC_K_UNET);
    k_conv_in<<<ew(d1.v), 256, 0, ctx->stream>>>(d1, u->x, P + L[L_IN].woff, P + L[L_IN].boff,
                                                 u->a1a, C1);
    conv(ctx, d1, u->a1a, C1, L[L_C1B].wb, P + L[L_C1B].boff, u->s1, C1, 27, 1, nullptr, 0);
    k_pool<<<ew(d2.v * C1), 256, 0, ctx->stream>>>(d2, u->s1, u->p1, C1);
    conv(ctx, d2, u->p1, C1, L[L_C2A].wb, P + L[L_C2A].boff, u->a2a, C2, 27, 1, nullptr, 0);
    conv(ctx, d2, u->a2a, C2, L[L_C2B].wb, P + L[L_C2B].boff, u->s2, C2, 27, 1, nullptr, 0);
    k_pool<<<ew(d3.v * C2), 256, 0, ctx->stream>>>(d3, u->s2, u->p2, C2);
    conv(ctx, d3, u->p2, C2, L[L_C3A].wb, P + L[L_C3A].boff, u->a3a, C3, 27, 1, nullptr, 0);
    conv(ctx, d3, u->a3a, C3, L[L_C3B].wb, P + L[L_C3B].boff, u->a3b, C3, 27, 1, nullptr, 0);
    conv(ctx, d3, u->a3b, C3, L[L_U2].wb, P + L[L_U2].boff, u->u2, C2, 1, 0, nullptr, 0);
    k_up_add<<<ew(d2.v * C2), 256, 0, ctx->stream>>>(d2, u->u2, u->s2, u->a2, C2);
    conv(ctx, d2, u->a2, C2, L[L_C2C].wb, P + L[L_C2C].boff, u->c2c, C2, 27, 1, nullptr, 0);
    conv(ctx, d2, u->c2c, C2, L[L_U1].wb, P + L[L_U1].boff, u->u1, C1, 1, 0, nullptr, 0);
    k_up_add<<<ew(d1.v * C1), 256, 0, ctx->stream>>>(d1, u->u1, u->s1, u->a1, C1);
    conv(ctx, d1, u->a1, C1, L[L_C1C].wb, P + L[L_C1C].boff, u->c1c, C1, 27, 1, nullptr, 0);
    k_conv_out<<<ew(d1.v), 256, 0, ctx->stream>>>(d1, u->c1c, C1, P + L[L_OUT].woff,
                                                  P + L[L_OUT].boff, u->y);
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

// backward pass from u->dy (FP32 [V1]); accumulates into u->g
static int unet_backward(pic_ctx* ctx, UNet* u) {
    const Dims &d1 = u->d1, &d2 = u->d2, &d3 = u->d3;
    const int C1 = u->C1, C2 = u->C2, C3 = u->C3;
    auto& L = u->L;
    float* G = u->g;
    PicProfScope ps(ctx, PIC_K_UNET);
    // output layer
    k_wgrad_out<<<grid_for(d1.v, DCHUNK), 256, 0, ctx->stream>>>(d1, u->c1c, C1, u->dy,
                                                                 G + L[L_OUT].woff, G + L[L_OUT].boff);
    k_conv_out_bwd<<<ew(d1.v), 256, 0, ctx->stream>>>(d1, u->dy, C1, u->p + L[L_OUT].woff, u->c1c,
                                                      u->gc1c);
    // c1c (input a1, no ReLU before it)
    wgrad(ctx, d1, u->a1, C1, u->gc1c, C1, 27, G + L[L_C1C].woff, G + L[L_C1C].boff);
    conv(ctx, d1, u->gc1c, C1, L[L_C1C].wt, nullptr, u->g1, C1, 27, 0, nullptr, 0);
    // a1 = up(u1) + s1: g(s1) starts as g1; g(u1) = sum over children
    k_up_bwd<<<ew(d2.v * C1), 256, 0, ctx->stream>>>(d2, u->g1, u->gu1, C1);
    // u1 (1x1, input c2c which is post-ReLU)
    wgrad(ctx, d2, u->c2c, C2, u->gu1, C1, 1, G + L[L_U1].woff, G + L[L_U1].boff);
    conv(ctx, d2, u->gu1, C1, L[L_U1].wt, nullptr, u->gc2c, C2, 1, 0, u->c2c, 0);
    // c2c (input a2)
    wgrad(ctx, d2, u->a2, C2, u->gc2c, C2, 27, G + L[L_C2C].woff, G + L[L_C2C].boff);
    conv(ctx, d2, u->gc2c, C2, L[L_C2C].wt, nullptr, u->g2, C2, 27, 0, nullptr, 0);
    // a2 = up(u2) + s2
    k_up_bwd<<<ew(d3.v * C2), 256, 0, ctx->stream>>>(d3, u->g2, u->gu2, C2);
    wgrad(ctx, d3, u->a3b, C3, u->gu2, C2, 1, G + L[L_U2].woff, G + L[L_U2].boff);
    conv(ctx, d3, u->gu2, C2, L[L_U2].wt, nullptr, u->g3b, C3, 1, 0, u->a3b, 0);
    // c3b (input a3a)
    wgrad(ctx, d3, u->a3a, C3, u->g3b, C3, 27, G + L[L_C3B].woff, G + L[L_C3B].boff);
    conv(ctx, d3, u->g3b, C3, L[L_C3B].wt, nullptr, u->g3a, C3, 27, 0, u->a3a, 0);
    // c3a (input p2 = pool(s2))
    wgrad(ctx, d3, u->p2, C2, u->g3a, C3, 27, G + L[L_C3A].woff, G + L[L_C3A].boff);
    conv(ctx, d3, u->g3a, C3, L[L_C3A].wt, nullptr, u->gp2, C2, 27, 0, nullptr, 0);
    // g(s2) = (g2 + pool_bwd(gp2)) * relu'(s2)   (g2 buffer reused in place)
    k_pool_bwd<<<ew(d2.v * C2), 256, 0, ctx->stream>>>(d2, u->gp2, u->g2, u->s2, C2);
    // c2b (input a2a)
    wgrad(ctx, d2, u->a2a, C2, u->g2, C2, 27, G + L[L_C2B].woff, G + L[L_C2B].boff);
    conv(ctx, d2, u->g2, C2, L[L_C2B].wt, nullptr, u->g2a, C2, 27, 0, u->a2a, 0);
    // c2a (input p1 = pool(s1))
    wgrad(ctx, d2, u->p1, C1, u->g2a, C2, 27, G + L[L_C2A].woff, G + L[L_C2A].boff);
    conv(ctx, d2, u->g2a, C2, L[L_C2A].wt, nullptr, u->gp1, C1, 27, 0, nullptr, 0);
    k_pool_bwd<<<ew(d1.v * C1), 256, 0, ctx->stream>>>(d1, u->gp1, u->g1, u->s1, C1);
    // c1b (input a1a)
    wgrad(ctx, d1, u->a1a, C1, u->g1, C1, 27, G + L[L_C1B].woff, G + L[L_C1B].boff);
    conv(ctx, d1, u->g1, C1, L[L_C1B].wt, nullptr, u->g1a, C1, 27, 0, u->a1a, 0);
    // input layer
    k_wgrad_in<<<grid_for(d1.v, DCHUNK), 256, 0, ctx->stream>>>(d1, u->x, u->g1a, C1,
                                                                G + L[L_IN].woff, G + L[L_IN].boff);
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

static UNet* get_unet(pic_ctx* ctx) { return ctx ? (UNet*)ctx->unet : nullptr; }

// ---------------------------------------------------------------- ABI
extern "C" int pic_unet_create(pic_ctx* ctx, const pic_unet_desc* desc) {
    if (!ctx || !desc) return PIC_ERR_ARG;
    int rc = unet_build(ctx, desc->c1, desc->c2, desc->c3, desc->seed);
    if (rc) return rc;
    return unet_sync_weights(ctx, get_unet(ctx));
}

extern "C" int pic_unet_param_count(pic_ctx* ctx, size_t* n) {
    UNet* u = get_unet(ctx);
    if (!u || !n) return PIC_ERR_STATE;
    *n = u->nparam;
    return PIC_OK;
}

extern "C" int pic_unet_get_params(pic_ctx* ctx, float* dst, size_t n) {
    UNet* u = get_unet(ctx);
    if (!u || !dst || n != u->nparam) return PIC_ERR_ARG;
    PIC_HIP(ctx, hipMemcpyAsync(dst, u->p, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

extern "C" int pic_unet_set_params(pic_ctx* ctx, const float* src, size_t n) {
    UNet* u = get_unet(ctx);
    if (!u || !src || n != u->nparam) return PIC_ERR_ARG;
    PIC_HIP(ctx, hipMemcpyAsync(u->p, src, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    int rc = unet_sync_weights(ctx, u);
    if (rc) return rc;
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    u->trained = true;
    return PIC_OK;
}

extern "C" int pic_unet_forward(pic_ctx* ctx, const float* in2, float* out) {
    UNet* u = get_unet(ctx);
    if (!u || !in2 || !out) return PIC_ERR_ARG;
    size_t V = (size_t)u->d1.v;
    PIC_HIP(ctx, hipMemcpyAsync(u->x, in2, 2 * V * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    int rc = unet_forward(ctx, u);
    if (rc) return rc;
    PIC_HIP(ctx, hipMemcpyAsync(out, u->y, V * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

extern "C" int pic_unet_loss_grad(pic_ctx* ctx, const float* in2, const float* target,
                                  double* loss, float* grad) {
    UNet* u = get_unet(ctx);
    if (!u || !in2 || !target) return PIC_ERR_ARG;
    size_t V = (size_t)u->d1.v;
    float* tdev = nullptr;
    PIC_HIP(ctx, hipMalloc(&tdev, V * sizeof(float)));
    hipError_t e = hipMemcpyAsync(tdev, target, V * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(u->x, in2, 2 * V * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(tdev);
        return pic_hip_check(ctx, e, "U-Net inputs");
    }
    size_t nvalid = 0;
    for (size_t q = 0; q < V; q++) nvalid += in2[2 * q + 1] <= 0.5f;
    int rc = unet_forward(ctx, u);
    if (!rc) {
        (void)hipMemsetAsync(ctx->scratch, 0, sizeof(double), ctx->stream);
        (void)hipMemsetAsync(u->g, 0, u->nparam * sizeof(float), ctx->stream);
        k_loss<<<ew(u->d1.v), 256, 0, ctx->stream>>>(u->d1.v, u->y, tdev, u->x,
                                                     1.0f / (float)std::max<size_t>(1, nvalid),
                                                     u->dy, ctx->scratch);
        rc = unet_backward(ctx, u);
    }
    double l = 0;
    if (!rc) {
        (void)hipMemcpyAsync(&l, ctx->scratch, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (grad)
            (void)hipMemcpyAsync(grad, u->g, u->nparam * sizeof(float), hipMemcpyDeviceToHost,
                                 ctx->stream);
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = pic_hip_check(ctx, e, "U-Net loss/grad");
    }
    (void)hipFree(tdev);
    if (loss) *loss = l;
    return rc;
}

// Adam step on the accumulated gradient (gscale multiplies the gradient)
static int adam_step(pic_ctx* ctx, UNet* u, float lr, float gscale) {
    u->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.f - std::pow(b1, (float)u->adam_t), c2 = 1.f - std::pow(b2, (float)u->adam_t);
    k_adam<<<ew((long long)u->nparam), 256, 0, ctx->stream>>>((long long)u->nparam, u->p, u->g, u->m,
                                                            u->v, lr, b1, b2, 1e-8f, c1, c2, gscale);
    PIC_HIP(ctx, hipGetLastError());
    return unet_sync_weights(ctx, u);
}

extern "C" int pic_unet_adam_step(pic_ctx* ctx, const float* grad, float lr) {
    UNet* u = get_unet(ctx);
    if (!u || !grad) return PIC_ERR_ARG;
    PIC_HIP(ctx, hipMemcpyAsync(u->g, grad, u->nparam * sizeof(float), hipMemcpyHostToDevice,
                                ctx->stream));
    int rc = adam_step(ctx, u, lr, 1.0f);
    if (rc) return rc;
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

// ------------------------------------------------------ data and preconditioner
namespace {

struct Blob {
    float cx, cy, cz, inv2s2, amp;
};

// Gaussian blobs plus optional Philox white noise; zero on Dirichlet nodes
__global__ void __launch_bounds__(256)
    k_blobs(Dims d, const Blob* __restrict__ blobs, int nb, float noise, uint32_t seed,
            uint32_t sample, const uint8_t* __restrict__ nt, float* __restrict__ out) {
    long long v = (long long)blockIdx.x * 256 + threadIdx.x;
    if (v >= d.v) return;
    if (nt[v] == NT_DIR) {
        out[v] = 0.f;
        return;
    }
    int xi, yi, zi;
    vox_xyz(d, v, xi, yi, zi);
    float s = 0.f;
    for (int b = 0; b < nb; b++) {
        float dx = xi - blobs[b].cx, dy = yi - blobs[b].cy, dz = zi - blobs[b].cz;
        s += blobs[b].amp * __expf(-(dx * dx + dy * dy + dz * dz) * blobs[b].inv2s2);
    }
    if (noise > 0.f) {
        uint32_t o[4];
        philox4x32((uint32_t)v, sample, 0x4E4F4953u, 0, seed, 0x55u, o);
        float u1 = (o[0] >> 8) * (1.0f / 16777216.0f) + 1e-7f, u2 = (o[1] >> 8) * (1.0f / 16777216.0f);
        s += noise * sqrtf(-2.f * __logf(u1)) * __cosf(6.2831853f * u2);
    }
    out[v] = s;
}

__global__ void __launch_bounds__(256)
    k_sumsq_f(long long n, const float* __restrict__ a, double* __restrict__ out) {
    double s = 0;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256)
        s += (double)a[e] * a[e];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
}

// in place: a *= 1/sqrt(sumsq/count); rho = eps0 * a / dx^2 (the plain solver's input)
__global__ void __launch_bounds__(256)
    k_norm_rho(long long n, float* __restrict__ a, const double* __restrict__ sumsq, double count,
               double rho_fac, double* __restrict__ rho) {
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    double s = sumsq[0] > 0 ? 1.0 / sqrt(sumsq[0] / count) : 1.0;
    float v = (float)(a[e] * s);
    a[e] = v;
    rho[e] = rho_fac * v;
}

__global__ void __launch_bounds__(256)
    k_d2f(long long n, const double* __restrict__ a, float* __restrict__ b) {
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) b[e] = (float)a[e];
}

__global__ void __launch_bounds__(256)
    k_make_input(long long n, const float* __restrict__ rt, const uint8_t* __restrict__ nt,
                 float* __restrict__ x2) {
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    x2[2 * e] = rt[e];
    x2[2 * e + 1] = nt[e] == NT_DIR ? 1.f : 0.f;
}

// preconditioner input: rho_tilde = -r dx^2 / s (unit RMS), s = rms(r dx^2)
__global__ void __launch_bounds__(256)
    k_input_from_residual(long long n, const double* __restrict__ r, const uint8_t* __restrict__ nt,
                          const double* __restrict__ sumsq, double dx2, double count,
                          float* __restrict__ x2) {
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    double s = sumsq[0] > 0 ? sqrt(sumsq[0] / count) : 1.0;  // rms of r (regular rows)
    bool dir = nt[e] == NT_DIR;
    x2[2 * e] = dir ? 0.f : (float)(-r[e] / s);
    x2[2 * e + 1] = dir ? 1.f : 0.f;
}

// A = L / dx^2, so A^-1 r = dx^2 L^-1 r; with rho_tilde = -r dx^2 / (s dx^2)
// the network returns phi_tilde = L^-1 (r / s), hence z = s dx^2 phi_tilde.
// (s is the RMS of r itself, so the unit-RMS input matches the training data.)
__global__ void __launch_bounds__(256)
    k_output_to_z(long long n, const float* __restrict__ y, const uint8_t* __restrict__ nt,
                  const double* __restrict__ sumsq, double dx2, double count, double* __restrict__ z) {
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    double s = sumsq[0] > 0 ? sqrt(sumsq[0] / count) : 1.0;
    z[e] = nt[e] == NT_DIR ? 0.0 : s * dx2 * (double)y[e];
}

__global__ void __launch_bounds__(256)
    k_sumsq_regular(long long n, const double* __restrict__ r, const uint8_t* __restrict__ nt,
                    double* __restrict__ out) {
    double s = 0;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256)
        if (nt[e] == NT_REG) s += r[e] * r[e];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
}

}  // namespace

static bool uniform_dx(const GridC& g) {
    return std::fabs(g.dx[1] - g.dx[0]) <= 1e-12 * g.dx[0] &&
           std::fabs(g.dx[2] - g.dx[0]) <= 1e-12 * g.dx[0];
}

bool pic_unet_ready(pic_ctx* ctx) {
    UNet* u = get_unet(ctx);
    return u && u->trained && uniform_dx(ctx->g);
}

// z = M(r): the network applied to the residual (device FP64 vectors)
int pic_unet_apply(pic_ctx* ctx, const double* r, double* z) {
    UNet* u = get_unet(ctx);
    if (!u) return pic_fail(ctx, PIC_ERR_STATE, "no U-Net");
    const GridC& g = ctx->g;
    long long n = g.nn;
    double dx2 = g.dx[0] * g.dx[0];
    double* ss = ctx->scratch + 64;
    PIC_HIP(ctx, hipMemsetAsync(ss, 0, sizeof(double), ctx->stream));
    unsigned nb = std::min<unsigned>(1024, ew(n));
    k_sumsq_regular<<<nb, 256, 0, ctx->stream>>>(n, r, ctx->ntype, ss);
    double count = (double)(g.ni - 2) * (g.nj - 2) * (g.nk - 2);
    k_input_from_residual<<<ew(n), 256, 0, ctx->stream>>>(n, r, ctx->ntype, ss, dx2, count, u->x);
    int rc = unet_forward(ctx, u);
    if (rc) return rc;
    k_output_to_z<<<ew(n), 256, 0, ctx->stream>>>(n, u->y, ctx->ntype, ss, dx2, count, z);
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

int pic_poisson_internal_sor(pic_ctx* ctx, double tol, int max_it, pic_solver_stats* st);

extern "C" int pic_unet_train(pic_ctx* ctx, const pic_unet_train_desc* td,
                              pic_unet_train_stats* stats) {
    UNet* u = get_unet(ctx);
    if (!u || !td) return pic_fail(ctx, PIC_ERR_STATE, "create the U-Net first");
    if (td->samples < 1 || td->epochs < 1 || td->batch < 1 || td->blobs_max < td->blobs_min)
        return pic_fail(ctx, PIC_ERR_ARG, "bad training parameters");
    const GridC& g = ctx->g;
    if (!uniform_dx(g)) return pic_fail(ctx, PIC_ERR_ARG, "the learned solver needs uniform dx");
    int rc = pic_ensure_ntype(ctx);
    if (rc) return rc;
    long long V = g.nn;
    size_t S = (size_t)td->samples;
    float *dens = nullptr, *targ = nullptr;
    double *phi_keep = nullptr, *rho_keep = nullptr;
    Blob* dblobs = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(dens);
        (void)hipFree(targ);
        (void)hipFree(phi_keep);
        (void)hipFree(rho_keep);
        (void)hipFree(dblobs);
    };
    hipError_t e = hipMalloc(&dens, S * V * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&targ, S * V * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&phi_keep, V * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&rho_keep, V * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dblobs, 64 * sizeof(Blob));
    if (e == hipSuccess) e = hipMemcpyAsync(phi_keep, ctx->phi, V * 8, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rho_keep, ctx->rho, V * 8, hipMemcpyDeviceToDevice, ctx->stream);
    if (e != hipSuccess) {
        cleanup();
        return pic_hip_check(ctx, e, "training buffers");
    }
    auto wall = []() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    };
    double t0 = wall();
    // ---- pairs from the plain solver ----
    uint32_t rs = td->seed * 747796405u + 2891336453u;
    auto urand = [&]() {
        rs ^= rs << 13;
        rs ^= rs >> 17;
        rs ^= rs << 5;
        return (rs >> 8) * (1.0 / 16777216.0);
    };
    double dx2 = g.dx[0] * g.dx[0];
    double count = (double)(g.ni - 2) * (g.nj - 2) * (g.nk - 2);
    double* ss = ctx->scratch + 64;
    for (size_t s = 0; s < S && !rc; s++) {
        int nb = td->blobs_min + (int)(urand() * (td->blobs_max - td->blobs_min + 1));
        nb = std::min(nb, 64);
        std::vector<Blob> hb(nb > 0 ? nb : 1);
        for (int b = 0; b < nb; b++) {
            double sig = td->sigma_min + urand() * (td->sigma_max - td->sigma_min);
            hb[b].cx = (float)(1 + urand() * (g.ni - 3));
            hb[b].cy = (float)(1 + urand() * (g.nj - 3));
            hb[b].cz = (float)(1 + urand() * (g.nk - 3));
            hb[b].inv2s2 = (float)(1.0 / (2 * sig * sig));
            hb[b].amp = (float)((urand() < 0.5 ? -1 : 1) * (0.3 + urand()));
        }
        float* dst = dens + s * V;
        e = hipMemcpyAsync(dblobs, hb.data(), hb.size() * sizeof(Blob), hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) {
            rc = pic_hip_check(ctx, e, "blobs");
            break;
        }
        k_blobs<<<ew(V), 256, 0, ctx->stream>>>(u->d1, dblobs, nb, (float)td->noise, td->seed,
                                               (uint32_t)s, ctx->ntype, dst);
        (void)hipMemsetAsync(ss, 0, sizeof(double), ctx->stream);
        k_sumsq_f<<<std::min<unsigned>(1024, ew(V)), 256, 0, ctx->stream>>>(V, dst, ss);
        k_norm_rho<<<ew(V), 256, 0, ctx->stream>>>(V, dst, ss, count, EPS0 / dx2, ctx->rho);
        (void)hipMemsetAsync(ctx->phi, 0, V * 8, ctx->stream);
        pic_solver_stats st;
        int r = pic_poisson_internal_sor(ctx, 1e-6, 20000, &st);
        if (r < 0) {
            rc = r;
            break;
        }
        k_d2f<<<ew(V), 256, 0, ctx->stream>>>(V, ctx->phi, targ + s * V);
    }
    (void)hipMemcpyAsync(ctx->phi, phi_keep, V * 8, hipMemcpyDeviceToDevice, ctx->stream);
    (void)hipMemcpyAsync(ctx->rho, rho_keep, V * 8, hipMemcpyDeviceToDevice, ctx->stream);
    e = hipStreamSynchronize(ctx->stream);
    if (!rc && e != hipSuccess) rc = pic_hip_check(ctx, e, "training data");
    double t1 = wall();
    // ---- Adam over shuffled samples ----
    std::vector<size_t> order(S);
    for (size_t s = 0; s < S; s++) order[s] = s;
    double first = 0, last = 0;
    int steps = 0;
    for (int ep = 0; ep < td->epochs && !rc; ep++) {
        for (size_t s = S - 1; s > 0; s--) std::swap(order[s], order[(size_t)(urand() * (s + 1)) % (s + 1)]);
        (void)hipMemsetAsync(ctx->scratch, 0, sizeof(double), ctx->stream);
        for (size_t b0 = 0; b0 < S && !rc; b0 += td->batch) {
            (void)hipMemsetAsync(u->g, 0, u->nparam * sizeof(float), ctx->stream);
            size_t b1 = std::min(S, b0 + (size_t)td->batch);
            for (size_t q = b0; q < b1 && !rc; q++) {
                size_t s = order[q];
                k_make_input<<<ew(V), 256, 0, ctx->stream>>>(V, dens + s * V, ctx->ntype, u->x);
                rc = unet_forward(ctx, u);
                if (rc) break;
                k_loss<<<ew(V), 256, 0, ctx->stream>>>(V, u->y, targ + s * V, u->x, (float)(1.0 / count),
                                                       u->dy, ctx->scratch);
                rc = unet_backward(ctx, u);
            }
            if (!rc) rc = adam_step(ctx, u, (float)td->lr, 1.0f / (float)(b1 - b0));
            steps++;
        }
        double l = 0;
        (void)hipMemcpyAsync(&l, ctx->scratch, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess && !rc) rc = pic_hip_check(ctx, e, "training");
        l /= (double)S;
        if (ep == 0) first = l;
        last = l;
    }
    double t2 = wall();
    cleanup();
    if (rc) return rc;
    u->trained = true;
    if (stats) {
        stats->loss_first = first;
        stats->loss_last = last;
        stats->seconds_data = t1 - t0;
        stats->seconds_train = t2 - t1;
        stats->steps = steps;
    }
    return PIC_OK;
}
