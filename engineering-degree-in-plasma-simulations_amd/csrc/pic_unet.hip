// pic_unet.hip -- learned warm start and preconditioner for the Poisson solve
// (no counterpart in the reference).
//
// Network: a D-level (2..5) 3-D convolutional U-Net.  Activations are
// channels-last BF16 ([voxel][C]); weights are BF16 for the MFMA layers with
// FP32 master copies; every accumulation is FP32.  With C_l channels on level
// l (grid halved per level):
//
//   level 0     a0 = relu(conv3(x, 2->C0))            direct kernel (FP32 input)
//               s0 = relu(conv3(a0, C0->C0))          MFMA
//   level l>0   p_l = pool(s_{l-1})
//               a_l = relu(conv3(p_l, C_{l-1}->C_l))  MFMA
//               s_l = relu(conv3(a_l, C_l->C_l))      MFMA
//   decoder     u_l = conv1(h_{l+1} or s_{D-1}, C_{l+1}->C_l)   MFMA (1x1)
//   l = D-2..0  m_l = up(u_l) + s_l                   nearest upsample + skip
//               h_l = relu(conv3(m_l, C_l->C_l))      MFMA
//   output      y = conv3(h_0, C0->1)                 direct kernel (FP32 output)
//
// Parameter order: in, B0, A1, B1, ..., A_{D-1}, B_{D-1}, U_{D-2}, Dec_{D-2},
// ..., U_0, Dec_0, out; each layer [cout][taps * cin] (k = tap * cin + ci)
// then [cout] biases.
//
// The network runs on the node grid coarsened by 1 or 2 (2x2x2 aggregates).
// Input channels: a charge density in grid units (normalised to unit RMS)
// and the fraction of non-regular (Dirichlet) nodes in the aggregate.
// Target: the mean over the aggregate's regular nodes of the potential of
// the unit-spacing problem laplacian(phi) = -rho, phi = 0 on Dirichlet nodes,
// divided by the same normalisation.  As a preconditioner the network is the
// coarse solve of a two-level cycle: the fine residual after smoothing is
// restricted to the aggregates, the network's correction is interpolated
// back trilinearly (cell-centred weights 3/4, 1/4).
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
#include <type_traits>
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
__device__ inline uint32_t pack2(float a, float b) {
    return (uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16);
}
__device__ inline void unpack8(const uint4& q, float* f) {
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        f[2 * k] = bf2f((bf16)(w[k] & 0xFFFFu));
        f[2 * k + 1] = bf2f((bf16)(w[k] >> 16));
    }
}

struct Dims {
    int nx, ny, nz;
    long long v;
};
Dims scaled_dims(const Dims& d, int f) {
    Dims h{d.nx / f, d.ny / f, d.nz / f, 0};
    h.v = (long long)h.nx * h.ny * h.nz;
    return h;
}

// 32-bit arithmetic: grids hold fewer than 2^31 voxels
__device__ inline void vox_xyz(const Dims& d, long long v, int& x, int& y, int& z) {
    unsigned vv = (unsigned)v, nx = (unsigned)d.nx, ny = (unsigned)d.ny;
    unsigned r = vv / nx;
    x = (int)(vv - r * nx);
    unsigned zz = r / ny;
    y = (int)(r - zz * ny);
    z = (int)zz;
}

// ---------------------------------------------------------------- MFMA conv
constexpr int TX = 16, TY = 4, TZ = 4;

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
    bf16* pool = nullptr;  // k_conv3_rb only: also the 2x2x2 mean of the output on the half grid
};

// NB N-blocks of 16 output channels per workgroup (cout/16 divisible by NB).
// Per 16-channel input chunk the workgroup stages the input tile (+ halo) and
// the chunk's weights for its NB*16 outputs in LDS; the tap loop then issues
// only LDS reads and MFMAs.
template <int TAPS, int NB>
__global__ void __launch_bounds__(256) k_conv_mfma(ConvArgs a) {
    constexpr int H = TAPS == 27 ? 1 : 0;
    constexpr int SX = TX + 2 * H, SY = TY + 2 * H, SZ = TZ + 2 * H;
    __shared__ uint4 tile[SX * SY * SZ * 2];  // 16 BF16 channels = 32 bytes per voxel
    __shared__ uint4 wl[NB * TAPS * 16 * 2];  // rows (nb*TAPS + t)*16 + n, 16 BF16 inputs each
    const Dims d = a.d;
    const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
    const int b = blockIdx.x;
    const int x0 = (b % ntx) * TX, y0 = ((b / ntx) % nty) * TY, z0 = (b / (ntx * nty)) * TZ;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, m = lane & 15;
    const int nb0 = blockIdx.y * NB;
    f4v acc[4][NB];
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[mb][nb] = f4v{0.f, 0.f, 0.f, 0.f};
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
        for (int s = tid; s < NB * TAPS * 16 * 2; s += 256) {
            int hf = s & 1, row = s >> 1;
            int n16 = row % 16, t = (row / 16) % TAPS, nb = row / (16 * TAPS);
            int n = (nb0 + nb) * 16 + n16;
            wl[s] = *reinterpret_cast<const uint4*>(a.w + (long long)n * TAPS * a.cin + t * a.cin + c0 +
                                                    hf * 8);
        }
        __syncthreads();
        for (int t = 0; t < TAPS; t++) {
            const int ox = TAPS == 27 ? t % 3 : 0, oy = TAPS == 27 ? (t / 3) % 3 : 0,
                      oz = TAPS == 27 ? t / 9 : 0;
            s4v bfr[NB];
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
                bfr[nb] = *reinterpret_cast<const s4v*>(reinterpret_cast<const char*>(wl) +
                                                        ((nb * TAPS + t) * 16 + m) * 32 + 8 * g);
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
                int row = wave * 4 + mb;
                int ty = row % TY, tz = row / TY;
                int sv = ((tz + oz) * SY + (ty + oy)) * SX + (m + ox);
                s4v afr = *reinterpret_cast<const s4v*>(reinterpret_cast<const char*>(tile) +
                                                        sv * 32 + 8 * g);
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
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
        for (int nb = 0; nb < NB; nb++) {
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

// 3x3x3 convolution with v_mfma_f32_16x16x32_bf16: K = 32 covers two taps of
// one 16-channel chunk (lane group g: tap 2p + (g >> 1), channels 8 (g & 1)..+7
// for both operands).  The next chunk's input tile and weights are fetched
// into registers while the current chunk computes.  SPLIT 1: 16x4x4 voxels,
// each wave owns 4 rows of 16 voxels and all taps.  SPLIT 4 (small grids):
// 16x2x2 voxels, every wave computes all 4 rows for every fourth tap pair and
// the partial sums are reduced through LDS.
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

template <int NB, int SPLIT, bool PF>
__global__ void __launch_bounds__(256) k_conv3(ConvArgs a) {
    constexpr int TY = SPLIT == 1 ? 4 : 2, TZ = TY;
    constexpr int SX = TX + 2, SY = TY + 2, SZ = TZ + 2;
    constexpr int NTILE = SX * SY * SZ * 2;  // uint4 per chunk (16 BF16 channels per voxel)
    constexpr int NW = NB * 27 * 16 * 2;     // uint4 per chunk: rows (nb*27 + t)*16 + n
    constexpr int RT = (NTILE + 255) / 256, RW = (NW + 255) / 256;
    constexpr int LDS_STAGE = (NTILE + NW) * 16;
    constexpr int LDS_RED = SPLIT == 1 ? 0 : 4 * 4 * NB * 64 * 16;
    constexpr int LDS = LDS_STAGE > LDS_RED ? LDS_STAGE : LDS_RED;
    __shared__ __attribute__((aligned(16))) unsigned char smem[LDS];
    uint4* tile = reinterpret_cast<uint4*>(smem);
    uint4* wl = tile + NTILE;
    const Dims d = a.d;
    const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
    const int b = blockIdx.x;
    const int x0 = (b % ntx) * TX, y0 = ((b / ntx) % nty) * TY, z0 = (b / (ntx * nty)) * TZ;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, m = lane & 15;
    const int hi = g >> 1, half = g & 1;
    const int nb0 = blockIdx.y * NB;
    const long long wstride = 27LL * a.cin;
    uint4 rt[PF ? RT : 1], rw[PF ? RW : 1];
    auto fetch = [&](int c0) {
        if (!PF) return;
#pragma unroll
        for (int q = 0; q < RT; q++) {
            int s = tid + 256 * q;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (s < NTILE) {
                int vox = s >> 1, hf = s & 1;
                int sx = vox % SX, sy = (vox / SX) % SY, sz = vox / (SX * SY);
                int gx = x0 + sx - 1, gy = y0 + sy - 1, gz = z0 + sz - 1;
                if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
                    v = *reinterpret_cast<const uint4*>(
                        a.in + (((long long)gz * d.ny + gy) * d.nx + gx) * a.cin + c0 + hf * 8);
            }
            rt[PF ? q : 0] = v;
        }
#pragma unroll
        for (int q = 0; q < RW; q++) {
            int s = tid + 256 * q;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (s < NW) {
                int hf = s & 1, row = s >> 1;
                int n16 = row % 16, t = (row / 16) % 27, nb = row / (16 * 27);
                int n = (nb0 + nb) * 16 + n16;
                v = *reinterpret_cast<const uint4*>(a.w + n * wstride + t * a.cin + c0 + hf * 8);
            }
            rw[PF ? q : 0] = v;
        }
    };
    f4v acc[4][NB];
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[mb][nb] = f4v{0.f, 0.f, 0.f, 0.f};
    // without prefetch: stage one element at a time (fewer registers)
    auto stage = [&](int c0) {
        for (int s = tid; s < NTILE; s += 256) {
            int vox = s >> 1, hf = s & 1;
            int sx = vox % SX, sy = (vox / SX) % SY, sz = vox / (SX * SY);
            int gx = x0 + sx - 1, gy = y0 + sy - 1, gz = z0 + sz - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
                v = *reinterpret_cast<const uint4*>(
                    a.in + (((long long)gz * d.ny + gy) * d.nx + gx) * a.cin + c0 + hf * 8);
            tile[s] = v;
        }
        for (int s = tid; s < NW; s += 256) {
            int hf = s & 1, row = s >> 1;
            int n16 = row % 16, t = (row / 16) % 27, nb = row / (16 * 27);
            int n = (nb0 + nb) * 16 + n16;
            wl[s] = *reinterpret_cast<const uint4*>(a.w + n * wstride + t * a.cin + c0 + hf * 8);
        }
    };
    if (PF) fetch(0);
    for (int c0 = 0; c0 < a.cin; c0 += 16) {
        __syncthreads();
        if (PF) {
#pragma unroll
            for (int q = 0; q < (PF ? RT : 1); q++)
                if (tid + 256 * q < NTILE) tile[tid + 256 * q] = rt[q];
#pragma unroll
            for (int q = 0; q < (PF ? RW : 1); q++)
                if (tid + 256 * q < NW) wl[tid + 256 * q] = rw[q];
        } else {
            stage(c0);
        }
        __syncthreads();
        if (PF && c0 + 16 < a.cin) fetch(c0 + 16);
        for (int p = (SPLIT == 1 ? 0 : wave); p < 14; p += SPLIT) {
            const int t = 2 * p + hi;
            const bool live = t < 27;  // the last pair has one tap; its second half is zero
            const int tt = live ? t : 26;
            const int ox = tt % 3, oy = (tt / 3) % 3, oz = tt / 9;
            bf8v bfr[NB];
#pragma unroll
            for (int nb = 0; nb < NB; nb++) {
                uint4 q = wl[((nb * 27 + tt) * 16 + m) * 2 + half];
                if (!live) q = make_uint4(0, 0, 0, 0);
                bfr[nb] = __builtin_bit_cast(bf8v, q);
            }
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
                int row = SPLIT == 1 ? wave * 4 + mb : mb;
                int ty = row % TY, tz = row / TY;
                int sv = ((tz + oz) * SY + (ty + oy)) * SX + (m + ox);
                uint4 q = tile[sv * 2 + half];
                if (!live) q = make_uint4(0, 0, 0, 0);
                bf8v afr = __builtin_bit_cast(bf8v, q);
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[nb], acc[mb][nb], 0, 0, 0);
            }
        }
    }
    constexpr int ROWS = SPLIT == 1 ? 4 : 1;
    if (SPLIT > 1) {
        __syncthreads();  // every wave is done with the staged chunk
        f4v* red = reinterpret_cast<f4v*>(smem);  // [wave][row][nb][lane]
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int nb = 0; nb < NB; nb++) red[((wave * 4 + mb) * NB + nb) * 64 + lane] = acc[mb][nb];
        __syncthreads();
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {  // wave w finalises row w
            f4v s4 = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int w2 = 0; w2 < 4; w2++) s4 += red[((w2 * 4 + wave) * NB + nb) * 64 + lane];
            acc[0][nb] = s4;
        }
    }
    // D fragment: lane holds rows 4g + j (x offset) of column m (channel)
#pragma unroll
    for (int mb = 0; mb < ROWS; mb++) {
        int row = SPLIT == 1 ? wave * 4 + mb : wave;
        int gy = y0 + row % TY, gz = z0 + row / TY;
        if (gy >= d.ny || gz >= d.nz) continue;
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
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
                a.out[o] = f2bf(v);
            }
        }
    }
}

// 3x3x3 convolution, register-blocked weights (large grids, 32 KS input
// channels).  A workgroup's 4 waves are NWN along the output channels (one
// 16-wide N-block each) times 4 / NWN along z; a wave computes 16 x RY x RZ
// output voxels (RY * RZ M-rows) of its N-block.  The input tile with a
// one-voxel halo and all channels is staged in LDS once (every global load
// issued before the first LDS store; 16-channel-chunk-major, 32 bytes per
// voxel).  Per 32-channel K step the wave keeps its N-block's 27 B fragments
// in registers and walks the input rows: the A fragment of input row
// (iy, iz) at x shift ox is read from LDS once and feeds the MFMAs of every
// output row (iy - dy, iz - dz) of the block, tap (ox, dy, dz).
template <int KS, int RY, int RZ, int NWN>
__global__ void __launch_bounds__(256, 2) k_conv3_rb(ConvArgs a) {  // two workgroups per CU
    constexpr int NWZ = 4 / NWN;
    constexpr int CIN = 32 * KS, NC = CIN / 16;
    constexpr int SX = TX + 2, SY = RY + 2, SZ = RZ * NWZ + 2, NV = SX * SY * SZ;
    constexpr int J = NC * 2;  // uint4 per voxel
    // chunk stride padded by 64 bytes: the staging stores (a voxel's J uint4
    // on consecutive lanes) fall on distinct banks
    constexpr int NVP = NV + 2;
    constexpr int NT = NV * J, RT = (NT + 255) / 256;
    __shared__ __attribute__((aligned(16))) uint4 tile[NVP * J];  // [chunk][voxel (NVP)][2]
    const Dims d = a.d;
    const int ntx = d.nx / TX, nty = d.ny / RY;
    const int b = blockIdx.x;
    const int x0 = (b % ntx) * TX, y0 = ((b / ntx) % nty) * RY, z0 = (b / (ntx * nty)) * (RZ * NWZ);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, m = lane & 15;
    const int wn = wave % NWN, wz = wave / NWN;
    const int n0 = (blockIdx.y * NWN + wn) * 16;
    const bf16* wrow = a.w + (long long)(n0 + m) * 27 * CIN + 8 * g;
    // the first K step's B fragments are requested before the input tile, so
    // the two loads' latencies overlap
    bf8v bfr[27];
#pragma unroll
    for (int t = 0; t < 27; t++)
        bfr[t] = __builtin_bit_cast(bf8v, *reinterpret_cast<const uint4*>(wrow + t * CIN));
    {
        // item e = voxel * J + j, j the 16-byte piece of the voxel's
        // channels: consecutive lanes read consecutive 16 bytes
        uint4 rt[RT];
#pragma unroll
        for (int q = 0; q < RT; q++) {
            const int e = tid + 256 * q;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (e < NT) {
                const int vox = e / J, j = e % J;
                const int sx = vox % SX, sy = (vox / SX) % SY, sz = vox / (SX * SY);
                const int gx = x0 + sx - 1, gy = y0 + sy - 1, gz = z0 + sz - 1;
                if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
                    v = *reinterpret_cast<const uint4*>(a.in + (((long long)gz * d.ny + gy) * d.nx + gx) * CIN + 8 * j);
            }
            rt[q] = v;
        }
#pragma unroll
        for (int q = 0; q < RT; q++) {
            const int e = tid + 256 * q;
            if (e < NT) tile[((e % J >> 1) * NVP + e / J) * 2 + (e & 1)] = rt[q];
        }
    }
    __syncthreads();
    f4v acc[RY][RZ];
#pragma unroll
    for (int ry = 0; ry < RY; ry++)
#pragma unroll
        for (int rz = 0; rz < RZ; rz++) acc[ry][rz] = f4v{0.f, 0.f, 0.f, 0.f};
    const int ch = g >> 1, hf = g & 1;  // lane group g: channels 8g..8g+7 of the K step
#pragma unroll
    for (int s = 0; s < KS; s++) {
        if (s > 0) {
#pragma unroll
            for (int t = 0; t < 27; t++)
                bfr[t] = __builtin_bit_cast(bf8v, *reinterpret_cast<const uint4*>(wrow + t * CIN + 32 * s));
        }
        const int c16 = 2 * s + ch;
        if constexpr (RY * RZ >= 16) {
            // 16 output rows per wave: registers leave no room for a queue
#pragma unroll
            for (int iz = 0; iz < RZ + 2; iz++)
#pragma unroll
                for (int iy = 0; iy < RY + 2; iy++)
#pragma unroll
                    for (int ox = 0; ox < 3; ox++) {
                        const int sv = ((wz * RZ + iz) * SY + iy) * SX + m + ox;
                        const bf8v afr = __builtin_bit_cast(bf8v, tile[(c16 * NVP + sv) * 2 + hf]);
#pragma unroll
                        for (int dz = 0; dz < 3; dz++)
#pragma unroll
                            for (int dy = 0; dy < 3; dy++) {
                                const int ry = iy - dy, rz = iz - dz;
                                if (ry < 0 || ry >= RY || rz < 0 || rz >= RZ) continue;
                                acc[ry][rz] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                    afr, bfr[(dz * 3 + dy) * 3 + ox], acc[ry][rz], 0, 0, 0);
                            }
                    }
        } else {
            // the (iz, iy, ox) walk flattened, A fragments read PD steps ahead
            // into a register queue so the LDS latency hides behind the MFMAs
            constexpr int NQ = (RZ + 2) * (RY + 2) * 3, PD = 4;
            auto afrag = [&](int q) {
                const int ox = q % 3, iy = (q / 3) % (RY + 2), iz = q / (3 * (RY + 2));
                const int sv = ((wz * RZ + iz) * SY + iy) * SX + m + ox;
                return __builtin_bit_cast(bf8v, tile[(c16 * NVP + sv) * 2 + hf]);
            };
            bf8v aq[PD];
#pragma unroll
            for (int q = 0; q < PD; q++) aq[q] = afrag(q);
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const bf8v afr = aq[q % PD];
                if (q + PD < NQ) aq[q % PD] = afrag(q + PD);
                __builtin_amdgcn_sched_barrier(0);  // keep the read ahead of these MFMAs
                const int ox = q % 3, iy = (q / 3) % (RY + 2), iz = q / (3 * (RY + 2));
#pragma unroll
                for (int dz = 0; dz < 3; dz++)
#pragma unroll
                    for (int dy = 0; dy < 3; dy++) {
                        const int ry = iy - dy, rz = iz - dz;
                        if (ry < 0 || ry >= RY || rz < 0 || rz >= RZ) continue;
                        acc[ry][rz] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            afr, bfr[(dz * 3 + dy) * 3 + ox], acc[ry][rz], 0, 0, 0);
                    }
            }
        }
    }
    // D fragment: lane holds rows 4g + j (x offset) of column m (channel).
    // The workgroup holds all CO = cout channels of its voxels (one N-block per
    // wave along N): the output tile is staged in LDS as BF16, channels-last,
    // then written with 16-byte stores (every (z, y) row of 16 voxels is
    // contiguous in memory).
    constexpr int CO = 16 * NWN, OZ = RZ * NWZ;
    static_assert(OZ * RY * TX * CO * 2 <= NVP * J * 16, "the output tile fits the staged input's LDS");
    __syncthreads();  // every wave is done with the staged input
    bf16* ot = reinterpret_cast<bf16*>(tile);
    const int n = n0 + m;
    const float bv = a.bias ? a.bias[n] : 0.f;
#pragma unroll
    for (int ry = 0; ry < RY; ry++)
#pragma unroll
        for (int rz = 0; rz < RZ; rz++) {
            const int gy = y0 + ry, gz = z0 + wz * RZ + rz;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                float v = acc[ry][rz][j] + bv;
                if (a.relu) v = fmaxf(v, 0.f);
                if (a.mask) {
                    const long long o = (((long long)gz * d.ny + gy) * d.nx + x0 + 4 * g + j) * a.cout + n;
                    if (!(bf2f(a.mask[o]) > 0.f)) v = 0.f;
                }
                ot[(((wz * RZ + rz) * RY + ry) * TX + 4 * g + j) * CO + wn * 16 + m] = f2bf(v);
            }
        }
    __syncthreads();
    {
        constexpr int NQ = OZ * RY * TX * CO / 8;  // uint4 in the output tile
        const uint4* oq = reinterpret_cast<const uint4*>(ot);
#pragma unroll
        for (int e = tid; e < NQ; e += 256) {
            const int c8 = e % (CO / 8), vox = e / (CO / 8);
            const int xx = vox % TX, yy = (vox / TX) % RY, zz = vox / (TX * RY);
            *reinterpret_cast<uint4*>(a.out + (((long long)(z0 + zz) * d.ny + y0 + yy) * d.nx + x0 + xx) * CO +
                                      8 * c8) = oq[e];
        }
    }
    if (a.pool) {
        // 2x2x2 means of the stored (BF16) outputs, children summed in
        // k_pool's order; every pooling cell lies in this lane's registers
        // (x pairs j, j+1; RY, RZ and the block origins are even)
        const int px0 = (x0 + 4 * g) / 2, hy = d.ny / 2, hx = d.nx / 2;
#pragma unroll
        for (int py = 0; py < RY / 2; py++)
#pragma unroll
            for (int pz = 0; pz < RZ / 2; pz++)
#pragma unroll
                for (int pxx = 0; pxx < 2; pxx++) {
                    float sum = 0.f;
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const int dx = q & 1, dy = (q >> 1) & 1, dz = q >> 2;
                        float v = acc[2 * py + dy][2 * pz + dz][2 * pxx + dx] + bv;
                        if (a.relu) v = fmaxf(v, 0.f);
                        sum += bf2f(f2bf(v));
                    }
                    const int cz = (z0 + wz * RZ) / 2 + pz, cy = y0 / 2 + py, cx = px0 + pxx;
                    a.pool[(((long long)cz * hy + cy) * hx + cx) * a.cout + n] = f2bf(0.125f * sum);
                }
    }
}

// 3x3x3 convolution for small grids (few workgroups): 16x2x2-voxel tiles,
// one 16-wide N-block per workgroup, all CIN input channels and the
// N-block's weights fetched in one round (every load issued before the first
// LDS store) and staged in LDS 16-channel-chunk-major (32 bytes per voxel or
// weight row, the conflict-free stride of the 16-byte fragment reads); K = 32
// channels of one tap per v_mfma_f32_16x16x32_bf16 step, the 27 * CIN / 32
// steps dealt over the 4 waves and reduced through LDS.
template <int CIN>
__global__ void __launch_bounds__(256) k_conv3_small(ConvArgs a) {
    constexpr int TY = 2, TZ = 2, SX = TX + 2, SY = TY + 2, SZ = TZ + 2;
    constexpr int NV = SX * SY * SZ;             // staged voxels
    constexpr int NC = CIN / 16;                 // 16-channel chunks
    constexpr int J = NC * 2;                    // uint4 per voxel or weight row
    // chunk strides padded by 64 bytes: the staging stores (a voxel's or
    // row's J uint4 on consecutive lanes) then fall on distinct banks
    constexpr int NVP = NV + 2, WR = 27 * 16, WRP = WR + 2;
    constexpr int NT = NV * J, NW = WR * J;      // uint4 to stage
    constexpr int RT = (NT + 255) / 256, RW = (NW + 255) / 256;
    constexpr int LDS_ST = (NVP + WRP) * J * 16, LDS_RED = 4 * 4 * 64 * 16;
    constexpr int LDS = LDS_ST > LDS_RED ? LDS_ST : LDS_RED;
    constexpr int KS = CIN / 32;                 // K steps per tap
    __shared__ __attribute__((aligned(16))) unsigned char smem[LDS];
    uint4* tile = reinterpret_cast<uint4*>(smem);  // [chunk][voxel (NVP)][2]
    uint4* wl = tile + NVP * J;                    // [chunk][tap * 16 + n (WRP)][2]
    const Dims d = a.d;
    const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
    const int b = blockIdx.x;
    const int x0 = (b % ntx) * TX, y0 = ((b / ntx) % nty) * TY, z0 = (b / (ntx * nty)) * TZ;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, m = lane & 15;
    const int nb0 = blockIdx.y;
    // staging: item e = (voxel or row) * J + j, j the 16-byte piece of its
    // channels, so consecutive lanes read consecutive 16 bytes
    uint4 rt[RT], rw[RW];
#pragma unroll
    for (int q = 0; q < RT; q++) {
        const int e = tid + 256 * q;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (e < NT) {
            const int vox = e / J, j = e % J;
            const int sx = vox % SX, sy = (vox / SX) % SY, sz = vox / (SX * SY);
            const int gx = x0 + sx - 1, gy = y0 + sy - 1, gz = z0 + sz - 1;
            if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
                v = *reinterpret_cast<const uint4*>(a.in + (((long long)gz * d.ny + gy) * d.nx + gx) * CIN + 8 * j);
        }
        rt[q] = v;
    }
    const bf16* wblk = a.w + (long long)nb0 * 16 * 27 * CIN;
#pragma unroll
    for (int q = 0; q < RW; q++) {
        const int e = tid + 256 * q;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (e < NW) {
            const int row = e / J, j = e % J;  // row = tap * 16 + n
            const int t = row >> 4, n = row & 15;
            v = *reinterpret_cast<const uint4*>(wblk + (n * 27 + t) * CIN + 8 * j);
        }
        rw[q] = v;
    }
#pragma unroll
    for (int q = 0; q < RT; q++) {
        const int e = tid + 256 * q;
        if (e < NT) tile[((e % J >> 1) * NVP + e / J) * 2 + (e & 1)] = rt[q];
    }
#pragma unroll
    for (int q = 0; q < RW; q++) {
        const int e = tid + 256 * q;
        if (e < NW) wl[((e % J >> 1) * WRP + e / J) * 2 + (e & 1)] = rw[q];
    }
    __syncthreads();
    f4v acc[4];
#pragma unroll
    for (int r = 0; r < 4; r++) acc[r] = f4v{0.f, 0.f, 0.f, 0.f};
    const int ch = g >> 1, hf = g & 1;  // lane group g: channels 8g..8g+7 of the K step
    for (int st = wave; st < 27 * KS; st += 4) {
        const int t = st / KS, h = st - t * KS;
        const int ox = t % 3, oy = (t / 3) % 3, oz = t / 9;
        const int c16 = 2 * h + ch;
        const bf8v bfr = __builtin_bit_cast(bf8v, wl[(c16 * WRP + t * 16 + m) * 2 + hf]);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ty = r % TY, tz = r / TY;
            const int sv = ((tz + oz) * SY + (ty + oy)) * SX + (m + ox);
            const bf8v afr = __builtin_bit_cast(bf8v, tile[(c16 * NVP + sv) * 2 + hf]);
            acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr, acc[r], 0, 0, 0);
        }
    }
    __syncthreads();  // every wave is done with the staged data
    f4v* red = reinterpret_cast<f4v*>(smem);  // [wave][row][lane]
#pragma unroll
    for (int r = 0; r < 4; r++) red[(wave * 4 + r) * 64 + lane] = acc[r];
    __syncthreads();
    f4v s4 = f4v{0.f, 0.f, 0.f, 0.f};  // wave w finalises row w
#pragma unroll
    for (int w2 = 0; w2 < 4; w2++) s4 += red[(w2 * 4 + wave) * 64 + lane];
    const int gy = y0 + wave % TY, gz = z0 + wave / TY;
    const bool row_in = gy < d.ny && gz < d.nz;
    const int n = nb0 * 16 + m;
    const float bv = a.bias ? a.bias[n] : 0.f;
    // D fragment: lane holds rows 4g + j (x offset) of column m (channel)
    float4 stored = make_float4(0.f, 0.f, 0.f, 0.f);  // the BF16 outputs, as floats
#pragma unroll
    for (int j = 0; j < 4; j++) {
        int gx = x0 + 4 * g + j;
        if (!row_in || gx >= d.nx) continue;
        long long o = (((long long)gz * d.ny + gy) * d.nx + gx) * a.cout + n;
        float v = s4[j] + bv;
        if (a.relu) v = fmaxf(v, 0.f);
        if (a.mask && !(bf2f(a.mask[o]) > 0.f)) v = 0.f;
        const bf16 vb = f2bf(v);
        a.out[o] = vb;
        (&stored.x)[j] = bf2f(vb);
    }
    if (a.pool) {
        // the 2x2x2 means of the stored outputs (the caller asks only for
        // even extents and nx a multiple of 16): the tile's four rows are
        // one 2x2 of (y, z), each lane holds two x pairs; children summed in
        // k_pool's order (x, then y, then z)
        __syncthreads();  // the reduction buffer is free
        float4* pr = reinterpret_cast<float4*>(smem);  // [wave][lane]
        pr[wave * 64 + lane] = stored;
        __syncthreads();
        if (wave == 0) {
            const int hy = d.ny / 2, hx = d.nx / 2;
#pragma unroll
            for (int px = 0; px < 2; px++) {
                float sum = 0.f;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int dx = q & 1, dy = (q >> 1) & 1, dz = q >> 2;
                    const float4 t = pr[(dy + 2 * dz) * 64 + lane];
                    sum += (&t.x)[2 * px + dx];
                }
                const int cx = (x0 + 4 * g) / 2 + px, cy = y0 / 2, cz = z0 / 2;
                a.pool[(((long long)cz * hy + cy) * hx + cx) * a.cout + n] = f2bf(0.125f * sum);
            }
        }
    }
}

// Input layer (2 -> 16 NB channels, 3x3x3, ReLU) on v_mfma_f32_16x16x32_bf16
// with K = tap * 2 + channel (54 used of two K = 32 steps): lane group g of
// step s holds taps 16 s + 4 g .. + 3, both channels, so a 16x4x4-voxel tile
// needs 2 MFMAs per row and N-block instead of one per tap.  The FP32 input
// x [V][2] is read directly; channel 0 is scaled by 1 / RMS from the striped
// sums nrm = {sum x^2, count} (1 without nrm).  w: BF16 [16 NB][64] (k >= 54
// zero); out: BF16 [V][16 NB].
template <int NB>
__global__ void __launch_bounds__(256)
    k_conv_in(Dims d, const float* __restrict__ x, const double* __restrict__ nrm,
              const bf16* __restrict__ w, const float* __restrict__ bias, bf16* __restrict__ out) {
    constexpr int TY = 4, TZ = 4, SX = TX + 2, SY = TY + 2, SZ = TZ + 2, CO = 16 * NB;
    __shared__ uint32_t tile[SX * SY * SZ];         // one BF16 pair per voxel
    __shared__ uint4 ot[TX * TY * TZ * CO / 8];     // output tile, channels-last BF16
    float sc = 1.f;
    if (nrm) {
        const double s0 = sget(nrm), s1 = sget(nrm + SSIZE);  // whole wave
        if (s0 > 0 && s1 > 0) sc = (float)(1.0 / sqrt(s0 / s1));
    }
    const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
    const int b = blockIdx.x;
    const int x0 = (b % ntx) * TX, y0 = ((b / ntx) % nty) * TY, z0 = (b / (ntx * nty)) * TZ;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, m = lane & 15;
    constexpr int NS = (SX * SY * SZ + 255) / 256;
    float2 xin[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {  // every load before the first LDS store
        const int sidx = tid + 256 * i;
        const int sx = sidx % SX, sy = (sidx / SX) % SY, sz = sidx / (SX * SY);
        const int gx = x0 + sx - 1, gy = y0 + sy - 1, gz = z0 + sz - 1;
        xin[i] = make_float2(0.f, 0.f);
        if (sidx < SX * SY * SZ && gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
            xin[i] = *reinterpret_cast<const float2*>(x + 2 * (((long long)gz * d.ny + gy) * d.nx + gx));
    }
#pragma unroll
    for (int i = 0; i < NS; i++)
        if (tid + 256 * i < SX * SY * SZ) tile[tid + 256 * i] = pack2(xin[i].x * sc, xin[i].y);
    // this lane's tap offsets in the tile for both K steps (-1: past tap 26)
    int off[2][4];
#pragma unroll
    for (int st = 0; st < 2; st++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int t = 16 * st + 4 * g + q;
            off[st][q] = t < 27 ? ((t / 9) * SY + (t / 3) % 3) * SX + t % 3 : -1;
        }
    bf8v bfr[2][NB];
#pragma unroll
    for (int st = 0; st < 2; st++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
            bfr[st][nb] = __builtin_bit_cast(
                bf8v, *reinterpret_cast<const uint4*>(w + (nb * 16 + m) * 64 + 32 * st + 8 * g));
    __syncthreads();
    f4v acc[4][NB];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        const int row = wave * 4 + mb;
        const int base = ((row / TY) * SY + row % TY) * SX + m;
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[mb][nb] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int st = 0; st < 2; st++) {
            uint4 q;
            q.x = off[st][0] >= 0 ? tile[base + off[st][0]] : 0u;
            q.y = off[st][1] >= 0 ? tile[base + off[st][1]] : 0u;
            q.z = off[st][2] >= 0 ? tile[base + off[st][2]] : 0u;
            q.w = off[st][3] >= 0 ? tile[base + off[st][3]] : 0u;
            bf8v afr = __builtin_bit_cast(bf8v, q);
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[st][nb], acc[mb][nb], 0, 0, 0);
        }
    }
    // D fragment: lane holds rows 4g + j (x offset) of column m (channel);
    // staged in LDS, then written with 16-byte stores
    bf16* otb = reinterpret_cast<bf16*>(ot);
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        const int row = wave * 4 + mb;
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            const int n = nb * 16 + m;
            const float bv = bias[n];
#pragma unroll
            for (int j = 0; j < 4; j++) otb[(row * TX + 4 * g + j) * CO + n] = f2bf(fmaxf(acc[mb][nb][j] + bv, 0.f));
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = tid; e < TX * TY * TZ * CO / 8; e += 256) {
        const int c8 = e % (CO / 8), vox = e / (CO / 8);
        const int row = vox / TX, gx = x0 + vox % TX, gy = y0 + row % TY, gz = z0 + row / TY;
        if (gx < d.nx && gy < d.ny && gz < d.nz)
            *reinterpret_cast<uint4*>(out + (((long long)gz * d.ny + gy) * d.nx + gx) * CO + 8 * c8) = ot[e];
    }
}

// Output layer (C -> 1, 3x3x3, no ReLU): y[v] = b + sum_t sum_c W[t][c] h[v + off_t][c].
// Per 16x4x4-voxel tile the per-tap projections P[v][t] = sum_c W[t][c] h[v][c]
// of the staged voxels (tile + one-voxel halo, zero outside the grid) come
// from v_mfma_f32_16x16x32_bf16 with M = voxels, N = taps (two blocks of 16,
// taps past 26 zero), K = the C channels (C = 16: upper half zero); P goes to
// LDS and every thread sums its voxel's 27 entries.  w: FP32 [27 C]
// (k = t C + c), rounded to BF16 like the other layers' weights.
template <int C>
__global__ void __launch_bounds__(256)
    k_conv_out(Dims d, const bf16* __restrict__ h, const float* __restrict__ w,
               const float* __restrict__ bias, float* __restrict__ y) {
    constexpr int TY = 4, TZ = 4, SX = TX + 2, SY = TY + 2, SZ = TZ + 2;
    constexpr int NV = SX * SY * SZ, MB = (NV + 15) / 16, MPW = (MB + 3) / 4;
    constexpr int PS = 27;  // LDS row stride of P (floats; odd: conflict-free gather)
    __shared__ float P[MB * 16 * PS];
    const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY;
    const int b = blockIdx.x;
    const int x0 = (b % ntx) * TX, y0 = ((b / ntx) % nty) * TY, z0 = (b / (ntx * nty)) * TZ;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, m = lane & 15;
    // B fragments: column n = tap nb * 16 + m, K elements = channels 8g..8g+7
    bf8v bfr[2];
#pragma unroll
    for (int nb = 0; nb < 2; nb++) {
        const int t = nb * 16 + m;
        uint32_t q[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int c = 8 * g + 2 * e;
            const bool ok = t < 27 && c < C;
            q[e] = ok ? pack2(w[t * C + c], w[t * C + c + 1]) : 0u;
        }
        bfr[nb] = __builtin_bit_cast(bf8v, make_uint4(q[0], q[1], q[2], q[3]));
    }
    // every A fragment of this wave's M-blocks is loaded before the first MFMA
    uint4 qa[MPW];
#pragma unroll
    for (int i = 0; i < MPW; i++) {
        const int vox = (wave + 4 * i) * 16 + m;
        uint4 q = make_uint4(0, 0, 0, 0);
        if (vox < NV && 8 * g < C) {
            const int sx = vox % SX, sy = (vox / SX) % SY, sz = vox / (SX * SY);
            const int gx = x0 + sx - 1, gy = y0 + sy - 1, gz = z0 + sz - 1;
            if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
                q = *reinterpret_cast<const uint4*>(h + (((long long)gz * d.ny + gy) * d.nx + gx) * C + 8 * g);
        }
        qa[i] = q;
    }
#pragma unroll
    for (int i = 0; i < MPW; i++) {
        const int mb = wave + 4 * i;
        if (mb >= MB) break;
        const bf8v afr = __builtin_bit_cast(bf8v, qa[i]);
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            f4v acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[nb], f4v{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            const int t = nb * 16 + m;
            if (t < 27) {
                // D fragment: lane holds rows 4g + j (voxels) of column m (tap)
#pragma unroll
                for (int j = 0; j < 4; j++) P[(mb * 16 + 4 * g + j) * PS + t] = acc[j];
            }
        }
    }
    __syncthreads();
    const int tx = tid % TX, ty = (tid / TX) % TY, tz = tid / (TX * TY);
    const int gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
    if (gx >= d.nx || gy >= d.ny || gz >= d.nz) return;
    float acc = bias[0];
#pragma unroll
    for (int t = 0; t < 27; t++) {
        const int ox = t % 3, oy = (t / 3) % 3, oz = t / 9;
        acc += P[(((tz + oz) * SY + ty + oy) * SX + tx + ox) * PS + t];
    }
    y[((long long)gz * d.ny + gy) * d.nx + gx] = acc;
}

// BF16 MFMA copy of the input layer's weights: wi[c][k] = w[c][k] for
// k = 2 tap + channel < 54, 0 up to 64
__global__ void k_pad_in_w(const float* __restrict__ win, int C, bf16* __restrict__ wi) {
    int e = blockIdx.x * 256 + threadIdx.x;
    if (e < C * 64) {
        int c = e / 64, k = e - c * 64;
        wi[e] = k < 54 ? f2bf(win[c * 54 + k]) : (bf16)0;
    }
}

// ------------------------------------------------------- direct conv layers
// d(input of the output layer) = sum_t dy[v - off t] w[t][c], times the ReLU
// mask of that input
template <int C>
__global__ void __launch_bounds__(256)
    k_conv_out_bwd(Dims d, const float* __restrict__ dy, const float* __restrict__ w,
                   const bf16* __restrict__ act, bf16* __restrict__ dx) {
    __shared__ float ws[27 * C];
    for (int t = threadIdx.x; t < 27 * C; t += 256) ws[t] = w[t];
    __syncthreads();
    long long v = (long long)blockIdx.x * 256 + threadIdx.x;
    if (v >= d.v) return;
    int xi, yi, zi;
    vox_xyz(d, v, xi, yi, zi);
    float acc[C];
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = 0.f;
    for (int t = 0; t < 27; t++) {
        int gx = xi - (t % 3 - 1), gy = yi - ((t / 3) % 3 - 1), gz = zi - (t / 9 - 1);
        if (gx < 0 || gx >= d.nx || gy < 0 || gy >= d.ny || gz < 0 || gz >= d.nz) continue;
        float g = dy[((long long)gz * d.ny + gy) * d.nx + gx];
#pragma unroll
        for (int c = 0; c < C; c++) acc[c] += g * ws[t * C + c];
    }
    const uint4* ar = reinterpret_cast<const uint4*>(act + v * C);
    uint4* o = reinterpret_cast<uint4*>(dx + v * C);
#pragma unroll
    for (int q = 0; q < C / 8; q++) {
        float f[8];
        unpack8(ar[q], f);
        float* a8 = acc + 8 * q;
#pragma unroll
        for (int k = 0; k < 8; k++) a8[k] = f[k] > 0.f ? a8[k] : 0.f;
        o[q] = make_uint4(pack2(a8[0], a8[1]), pack2(a8[2], a8[3]), pack2(a8[4], a8[5]),
                          pack2(a8[6], a8[7]));
    }
}

// ------------------------------------------------------ pooling / upsampling
// one thread per (voxel, group of 8 channels); C is a multiple of 16
__device__ inline long long child(const Dims& dc, int xi, int yi, int zi, int q) {
    int fx = 2 * dc.nx, fy = 2 * dc.ny;
    return ((long long)(2 * zi + (q >> 2)) * fy + (2 * yi + ((q >> 1) & 1))) * fx + (2 * xi + (q & 1));
}
__device__ inline long long parent(const Dims& df, int xi, int yi, int zi) {
    return ((long long)(zi >> 1) * (df.ny >> 1) + (yi >> 1)) * (df.nx >> 1) + (xi >> 1);
}

__global__ void __launch_bounds__(256)
    k_pool(Dims dc, const bf16* __restrict__ fine, bf16* __restrict__ coarse, int C) {
    const int G = C / 8;
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= dc.v * G) return;
    long long v = e / G;
    int c8 = (int)(e - v * G);
    int xi, yi, zi;
    vox_xyz(dc, v, xi, yi, zi);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < 8; q++) {
        float f[8];
        unpack8(reinterpret_cast<const uint4*>(fine + child(dc, xi, yi, zi, q) * C)[c8], f);
#pragma unroll
        for (int k = 0; k < 8; k++) s[k] += f[k];
    }
    reinterpret_cast<uint4*>(coarse + v * C)[c8] =
        make_uint4(pack2(0.125f * s[0], 0.125f * s[1]), pack2(0.125f * s[2], 0.125f * s[3]),
                   pack2(0.125f * s[4], 0.125f * s[5]), pack2(0.125f * s[6], 0.125f * s[7]));
}

// d_fine = (d_fine + d_coarse[parent] / 8) * (act > 0)
__global__ void __launch_bounds__(256)
    k_pool_bwd(Dims df, const bf16* __restrict__ dcoarse, bf16* __restrict__ dfine,
               const bf16* __restrict__ act, int C) {
    const int G = C / 8;
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= df.v * G) return;
    long long v = e / G;
    int c8 = (int)(e - v * G);
    int xi, yi, zi;
    vox_xyz(df, v, xi, yi, zi);
    float a[8], b[8], m[8], r[8];
    unpack8(reinterpret_cast<const uint4*>(dfine + v * C)[c8], a);
    unpack8(reinterpret_cast<const uint4*>(dcoarse + parent(df, xi, yi, zi) * C)[c8], b);
    unpack8(reinterpret_cast<const uint4*>(act + v * C)[c8], m);
#pragma unroll
    for (int k = 0; k < 8; k++) r[k] = m[k] > 0.f ? a[k] + 0.125f * b[k] : 0.f;
    reinterpret_cast<uint4*>(dfine + v * C)[c8] =
        make_uint4(pack2(r[0], r[1]), pack2(r[2], r[3]), pack2(r[4], r[5]), pack2(r[6], r[7]));
}

// fine = up(coarse) + skip
__global__ void __launch_bounds__(256)
    k_up_add(Dims df, const bf16* __restrict__ coarse, const bf16* __restrict__ skip,
             bf16* __restrict__ fine, int C) {
    const int G = C / 8;
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= df.v * G) return;
    long long v = e / G;
    int c8 = (int)(e - v * G);
    int xi, yi, zi;
    vox_xyz(df, v, xi, yi, zi);
    float a[8], b[8], r[8];
    unpack8(reinterpret_cast<const uint4*>(coarse + parent(df, xi, yi, zi) * C)[c8], a);
    unpack8(reinterpret_cast<const uint4*>(skip + v * C)[c8], b);
#pragma unroll
    for (int k = 0; k < 8; k++) r[k] = a[k] + b[k];
    reinterpret_cast<uint4*>(fine + v * C)[c8] =
        make_uint4(pack2(r[0], r[1]), pack2(r[2], r[3]), pack2(r[4], r[5]), pack2(r[6], r[7]));
}

// Decoder input m = up(conv1(h)) + skip on the fine level: the 1x1
// up-convolution of the coarse activation h (32 KS channels) on
// v_mfma_f32_16x16x32_bf16 (16 coarse voxels per workgroup, wave w computes
// N-block w of the 16 NB output channels, bias added), staged in LDS, then
// nearest upsampling plus the skip connection for the 8 children of each
// coarse voxel, 16 bytes per access.  Fine dimensions are exactly twice the
// coarse ones.  w: BF16 [16 NB][32 KS].
template <int KS, int NB>
__global__ void __launch_bounds__(256)
    k_up1_add(Dims dc, const bf16* __restrict__ h, const bf16* __restrict__ w,
              const float* __restrict__ bias, const bf16* __restrict__ skip, bf16* __restrict__ out) {
    static_assert(NB <= 4, "one N-block per wave");
    constexpr int CN = 32 * KS, CL = 16 * NB, G = CL / 8, US = CL + 4, VB = 16;
    __shared__ float ul[VB * US];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, m = lane & 15;
    const long long v0 = (long long)blockIdx.x * VB;
    if (wave < NB) {
        f4v acc = f4v{0.f, 0.f, 0.f, 0.f};
        const long long va = v0 + m;  // A row m
#pragma unroll
        for (int st = 0; st < KS; st++) {
            uint4 q = make_uint4(0, 0, 0, 0);
            if (va < dc.v) q = *reinterpret_cast<const uint4*>(h + va * CN + 32 * st + 8 * g);
            const bf8v bfr = __builtin_bit_cast(
                bf8v, *reinterpret_cast<const uint4*>(w + (wave * 16 + m) * CN + 32 * st + 8 * g));
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, q), bfr, acc, 0, 0, 0);
        }
        // D fragment: lane holds rows 4g + j (coarse voxels) of column m (channel)
        const float bv = bias[wave * 16 + m];
#pragma unroll
        for (int j = 0; j < 4; j++) ul[(4 * g + j) * US + wave * 16 + m] = acc[j] + bv;
    }
    __syncthreads();
    const int fx = 2 * dc.nx, fy = 2 * dc.ny;
    // items ordered (y/z child, coarse voxel, x child, 8-channel group) so that
    // consecutive threads touch consecutive 16-byte pieces of a fine x-row;
    // every skip load is issued before the first store
    constexpr int NITEM = 4 * VB * 2 * G, NIT = (NITEM + 255) / 256;
    long long fo[NIT];
    uint4 sq[NIT];
#pragma unroll
    for (int i = 0; i < NIT; i++) {
        const int it = tid + 256 * i;
        const int c8 = it % G, ax = (it / G) & 1, vl = (it / (2 * G)) % VB, bc = it / (2 * G * VB);
        const long long v = v0 + vl;
        fo[i] = -1;
        sq[i] = make_uint4(0, 0, 0, 0);
        if (it < NITEM && v < dc.v) {
            int cx, cy, cz;
            vox_xyz(dc, v, cx, cy, cz);
            fo[i] = ((long long)(2 * cz + (bc >> 1)) * fy + 2 * cy + (bc & 1)) * fx + 2 * cx + ax;
            sq[i] = reinterpret_cast<const uint4*>(skip + fo[i] * CL)[c8];
        }
    }
#pragma unroll
    for (int i = 0; i < NIT; i++) {
        if (fo[i] < 0) continue;
        const int it = tid + 256 * i;
        const int c8 = it % G, vl = (it / (2 * G)) % VB;
        float sk[8];
        unpack8(sq[i], sk);
        const float* uu = ul + vl * US + 8 * c8;
        reinterpret_cast<uint4*>(out + fo[i] * CL)[c8] =
            make_uint4(pack2(uu[0] + sk[0], uu[1] + sk[1]), pack2(uu[2] + sk[2], uu[3] + sk[3]),
                       pack2(uu[4] + sk[4], uu[5] + sk[5]), pack2(uu[6] + sk[6], uu[7] + sk[7]));
    }
}

// d_coarse = sum over the 8 children of d_fine
__global__ void __launch_bounds__(256)
    k_up_bwd(Dims dc, const bf16* __restrict__ dfine, bf16* __restrict__ dcoarse, int C) {
    const int G = C / 8;
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= dc.v * G) return;
    long long v = e / G;
    int c8 = (int)(e - v * G);
    int xi, yi, zi;
    vox_xyz(dc, v, xi, yi, zi);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < 8; q++) {
        float f[8];
        unpack8(reinterpret_cast<const uint4*>(dfine + child(dc, xi, yi, zi, q) * C)[c8], f);
#pragma unroll
        for (int k = 0; k < 8; k++) s[k] += f[k];
    }
    reinterpret_cast<uint4*>(dcoarse + v * C)[c8] =
        make_uint4(pack2(s[0], s[1]), pack2(s[2], s[3]), pack2(s[4], s[5]), pack2(s[6], s[7]));
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
    const int cin8 = cin / 8, cout8 = cout / 8;
    for (int tt = 0; tt < tiles_per_block; tt++) {
        int tile = blockIdx.x * tiles_per_block + tt;
        if (tile >= ntiles) break;
        int x0 = (tile % ntx) * WX, y0 = ((tile / ntx) % nty) * WY, z0 = (tile / (ntx * nty)) * WZ;
        __syncthreads();
        for (int s = threadIdx.x; s < SX * SY * SZ * cin8; s += 256) {
            int vox = s / cin8, c8 = s - vox * cin8;
            int sx = vox % SX, sy = (vox / SX) % SY, sz = vox / (SX * SY);
            int gx = x0 + sx - H, gy = y0 + sy - H, gz = z0 + sz - H;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
                val = reinterpret_cast<const uint4*>(X + (((long long)gz * d.ny + gy) * d.nx + gx) * cin)[c8];
            reinterpret_cast<uint4*>(xs)[s] = val;
        }
        for (int s = threadIdx.x; s < WX * WY * WZ * cout8; s += 256) {
            int vox = s / cout8, c8 = s - vox * cout8;
            int sx = vox % WX, sy = (vox / WX) % WY, sz = vox / (WX * WY);
            int gx = x0 + sx, gy = y0 + sy, gz = z0 + sz;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (gx < d.nx && gy < d.ny && gz < d.nz)
                val = reinterpret_cast<const uint4*>(dY + (((long long)gz * d.ny + gy) * d.nx + gx) * cout)[c8];
            reinterpret_cast<uint4*>(ys)[s] = val;
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
                uint2 xq = *reinterpret_cast<const uint2*>(xr);
                uint2 yq = *reinterpret_cast<const uint2*>(yr);
                float xv[4] = {bf2f(xq.x & 0xFFFFu), bf2f(xq.x >> 16), bf2f(xq.y & 0xFFFFu),
                               bf2f(xq.y >> 16)};
                float yv[4] = {bf2f(yq.x & 0xFFFFu), bf2f(yq.x >> 16), bf2f(yq.y & 0xFFFFu),
                               bf2f(yq.y >> 16)};
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

template <typename T>
__device__ inline float to_f(T v);
template <>
__device__ inline float to_f<float>(float v) {
    return v;
}
template <>
__device__ inline float to_f<bf16>(bf16 v) {
    return bf2f(v);
}

// direct layers: dW[o][t*XC + i] += sum_v G[v][o] X[v + off t][i] and
// db[o] += sum_v G[v][o] (in layer: G = dY (C ch), X = x (2 ch); out layer:
// G = dy (1 ch), X = h0 (C ch)).  Same tiling as k_wgrad, one weight per
// thread slot, RPT slots per thread.
template <int GC, int XC, typename GT, typename XT>
__global__ void __launch_bounds__(256)
    k_wgrad_direct(Dims d, const XT* __restrict__ X, const GT* __restrict__ G,
                   float* __restrict__ dW, float* __restrict__ db, int tiles_per_block) {
    constexpr int SX = WX + 2, SY = WY + 2, SZ = WZ + 2;
    constexpr int PER_O = 27 * XC + 1;
    constexpr int NW = GC * PER_O;
    constexpr int RPT = (NW + 255) / 256;
    __shared__ float xs[SX * SY * SZ * XC];
    __shared__ float gs[WX * WY * WZ * GC];
    const int ntx = (d.nx + WX - 1) / WX, nty = (d.ny + WY - 1) / WY, ntz = (d.nz + WZ - 1) / WZ;
    const int ntiles = ntx * nty * ntz;
    float acc[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) acc[r] = 0.f;
    for (int tt = 0; tt < tiles_per_block; tt++) {
        int tile = blockIdx.x * tiles_per_block + tt;
        if (tile >= ntiles) break;
        int x0 = (tile % ntx) * WX, y0 = ((tile / ntx) % nty) * WY, z0 = (tile / (ntx * nty)) * WZ;
        __syncthreads();
        for (int s = threadIdx.x; s < SX * SY * SZ * XC; s += 256) {
            int vox = s / XC, c = s - vox * XC;
            int sx = vox % SX, sy = (vox / SX) % SY, sz = vox / (SX * SY);
            int gx = x0 + sx - 1, gy = y0 + sy - 1, gz = z0 + sz - 1;
            float val = 0.f;
            if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
                val = to_f<XT>(X[(((long long)gz * d.ny + gy) * d.nx + gx) * XC + c]);
            xs[s] = val;
        }
        for (int s = threadIdx.x; s < WX * WY * WZ * GC; s += 256) {
            int vox = s / GC, c = s - vox * GC;
            int sx = vox % WX, sy = (vox / WX) % WY, sz = vox / (WX * WY);
            int gx = x0 + sx, gy = y0 + sy, gz = z0 + sz;
            float val = 0.f;
            if (gx < d.nx && gy < d.ny && gz < d.nz)
                val = to_f<GT>(G[(((long long)gz * d.ny + gy) * d.nx + gx) * GC + c]);
            gs[s] = val;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            int wi = threadIdx.x + 256 * r;
            if (wi >= NW) continue;
            int o = wi / PER_O, k = wi - o * PER_O;
            float a = 0.f;
            if (k == 27 * XC) {
                for (int v = 0; v < WX * WY * WZ; v++) a += gs[v * GC + o];
            } else {
                int t = k / XC, i = k - t * XC;
                int ox = t % 3, oy = (t / 3) % 3, oz = t / 9;
                for (int v = 0; v < WX * WY * WZ; v++) {
                    int vx = v % WX, vy = (v / WX) % WY, vz = v / (WX * WY);
                    a += gs[v * GC + o] * xs[(((vz + oz) * SY + (vy + oy)) * SX + (vx + ox)) * XC + i];
                }
            }
            acc[r] += a;
        }
    }
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        int wi = threadIdx.x + 256 * r;
        if (wi >= NW) continue;
        int o = wi / PER_O, k = wi - o * PER_O;
        if (k == 27 * XC)
            atomicAdd(&db[o], acc[r]);
        else
            atomicAdd(&dW[o * 27 * XC + k], acc[r]);
    }
}

// bias gradient of an MFMA layer: db[c] += sum_v dY[v][c]; partial sums of a
// workgroup are combined in LDS, one atomic per channel and workgroup
__global__ void __launch_bounds__(256)
    k_bias_grad(long long V, const bf16* __restrict__ dY, int C, float* __restrict__ db) {
    __shared__ float red[256];
    int c = threadIdx.x % C;
    int per = 256 / C;  // voxel lanes
    int lane = threadIdx.x / C;
    float acc = 0.f;
    if (lane < per) {
        long long step = (long long)gridDim.x * per;
        for (long long v = (long long)blockIdx.x * per + lane; v < V; v += step) acc += bf2f(dY[v * C + c]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < C) {
        float s = 0.f;
        for (int l = 0; l < per; l++) s += red[l * C + threadIdx.x];
        atomicAdd(&db[threadIdx.x], s);
    }
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

// loss over voxels holding a regular node (mask channel < 1):
// dy = (y - t) / Nvalid, out[0] += 0.5 sum (y-t)^2 / Nvalid
__global__ void __launch_bounds__(256)
    k_loss(long long V, const float* __restrict__ y, const float* __restrict__ t,
           const float* __restrict__ x2, float inv_n, float* __restrict__ dy,
           double* __restrict__ out) {
    long long v = (long long)blockIdx.x * 256 + threadIdx.x;
    double l = 0;
    if (v < V) {
        float valid = x2[2 * v + 1] < 0.999f ? 1.f : 0.f;
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

template <int N>
__device__ inline void block_sums(double (&v)[N]) {
    __shared__ double red[N][4];
#pragma unroll
    for (int q = 0; q < N; q++) {
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < N; q++) v[q] = red[q][0] + red[q][1] + red[q][2] + red[q][3];
}

__device__ inline void block_sum2(double& a, double& b) {
    __shared__ double red[2][4];
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        b = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

// ------------------------------------------- fine grid <-> network grid
// The node (i, j, k) belongs to aggregate (i/CF, j/CF, k/CF).

// training target: tgt[v] = mean of phi over the aggregate's regular nodes;
// striped sums + 2 SSIZE += tgt^2 over aggregates holding a regular node
template <int CF>
__global__ void __launch_bounds__(256)
    k_restrict_target(Dims df, Dims dc, const uint8_t* __restrict__ nt,
                      const double* __restrict__ phi, float* __restrict__ tgt,
                      double* __restrict__ sums) {
    double acc = 0;
    long long end;
    for (long long v = xcd_chunk(dc.v, 256, end) + threadIdx.x; v < end; v += 256) {
        int cx, cy, cz;
        vox_xyz(dc, v, cx, cy, cz);
        double sp = 0;
        int nreg = 0;
#pragma unroll
        for (int q = 0; q < CF * CF * CF; q++) {
            int fx = CF * cx + q % CF, fy = CF * cy + (q / CF) % CF, fz = CF * cz + q / (CF * CF);
            long long u = ((long long)fz * df.ny + fy) * df.nx + fx;
            if (nt[u] == NT_REG) {
                sp += phi[u];
                nreg++;
            }
        }
        float tv = nreg ? (float)(sp / nreg) : 0.f;
        tgt[v] = tv;
        acc += (double)tv * tv;
    }
    double dummy = 0;
    block_sum2(acc, dummy);
    if (threadIdx.x == 0) sadd(sums + 2 * SSIZE, acc);
}

// input to unit RMS over the valid aggregates, the target by the same factor
// (the map stays linear in the input's scale)
__global__ void __launch_bounds__(256)
    k_scale_pair(long long n, const double* __restrict__ sums, float* __restrict__ x2,
                 float* __restrict__ tgt) {
    long long v = (long long)blockIdx.x * 256 + threadIdx.x;
    const double s0 = sget(sums), s1 = sget(sums + SSIZE);  // striped, whole wave
    if (v >= n) return;
    double sx = s0 > 0 && s1 > 0 ? sqrt(s0 / s1) : 1.0;
    float f = (float)(1.0 / sx);
    x2[2 * v] *= f;
    tgt[v] *= f;
}

__global__ void __launch_bounds__(256) k_scale_f(long long n, float f, float* __restrict__ a) {
    long long v = (long long)blockIdx.x * 256 + threadIdx.x;
    if (v < n) a[v] *= f;
}

// preconditioner input from the fine residual of A e = t after smoothing z:
// x2[v] = {-h^2 mean over regular nodes of (t - A z), fraction of other nodes};
// striped sums (sums, sums + SSIZE) += {x^2, 1}; res (optional) receives
// t - A z on the fine grid (FP32, 0 on non-regular nodes).  z is FP32.  Regular nodes are
// interior (Dirichlet walls), so all six neighbours exist; z is zero on
// non-regular nodes.  CF 2 loads z once per aggregate (the 2x2x2 cube and
// its 24 face neighbours, coordinates clamped into the grid; values of
// non-regular nodes are discarded), so the loop has no branches.
template <int CF, typename TT = double>
__global__ void __launch_bounds__(256)
    k_res_restrict(Dims df, Dims dc, double h2, const uint8_t* __restrict__ nt,
                   const TT* __restrict__ t, const float* __restrict__ z, float* __restrict__ x2,
                   double* __restrict__ sums, float* __restrict__ res, const double* __restrict__ P) {
    double s2 = 0, cnt = 0;
    const unsigned sj = df.nx, sk = (unsigned)df.nx * df.ny;
    const double ih2 = 1.0 / h2;
    long long end;
    for (long long v = xcd_chunk(dc.v, 256, end) + threadIdx.x; v < end; v += 256) {
        int cx, cy, cz;
        vox_xyz(dc, v, cx, cy, cz);
        double acc = 0;
        int nreg = 0;
        if constexpr (CF == 2) {
            // clamped fine coordinates -1..2 around the aggregate, per axis
            unsigned X[4], Y[4], Z[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                X[q] = (unsigned)min(max(2 * cx - 1 + q, 0), df.nx - 1);
                Y[q] = (unsigned)min(max(2 * cy - 1 + q, 0), df.ny - 1) * sj;
                Z[q] = (unsigned)min(max(2 * cz - 1 + q, 0), df.nz - 1) * sk;
            }
            auto zat = [&](int a, int b, int c) { return (double)z[Z[c] + Y[b] + X[a]]; };
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int a = 1 + (q & 1), b = 1 + ((q >> 1) & 1), c = 1 + (q >> 2);
                const unsigned u = Z[c] + Y[b] + X[a];
                const bool reg = nt[u] == NT_REG;
                double az = (zat(a - 1, b, c) + zat(a + 1, b, c) + zat(a, b - 1, c) + zat(a, b + 1, c) +
                             zat(a, b, c - 1) + zat(a, b, c + 1) - 6.0 * zat(a, b, c)) *
                            ih2;
                if (P) az -= P[u] * zat(a, b, c);  // J z, J = A - diag(P)
                double rv = reg ? (double)t[u] - az : 0.0;
                acc += rv;
                nreg += reg;
                if (res) res[u] = (float)rv;
            }
        } else {
#pragma unroll
            for (int q = 0; q < CF * CF * CF; q++) {
                unsigned fx = CF * cx + q % CF, fy = CF * cy + (q / CF) % CF, fz = CF * cz + q / (CF * CF);
                unsigned u = (fz * df.ny + fy) * df.nx + fx;
                const bool reg = nt[u] == NT_REG;
                const unsigned oi = reg ? 1u : 0u, oj = reg ? sj : 0u, ok = reg ? sk : 0u;
                double az = ((double)z[u - oi] + z[u + oi] + z[u - oj] + z[u + oj] + z[u - ok] + z[u + ok] -
                             6.0 * z[u]) *
                            ih2;
                if (P) az -= P[u] * (double)z[u];
                double rv = reg ? (double)t[u] - az : 0.0;
                acc += rv;
                nreg += reg;
                if (res) res[u] = (float)rv;
            }
        }
        float xv = nreg ? (float)(-h2 * acc / nreg) : 0.f;
        x2[2 * v] = xv;
        x2[2 * v + 1] = 1.f - (float)nreg / (float)(CF * CF * CF);
        if (nreg) {
            s2 += (double)xv * xv;
            cnt += 1;
        }
    }
    block_sum2(s2, cnt);
    if (threadIdx.x == 0) {
        sadd(sums, s2);
        sadd(sums + SSIZE, cnt);
    }
}


// k_res_restrict<2, float> with even fine extents (the fine grid exactly twice
// the network grid): each x-row of an aggregate's neighbourhood comes in
// 8-byte pieces (z at 2cx, 2cx+1 and the pieces on either side for the
// i-neighbours; r, the node types, P and res per row likewise); the same
// arithmetic per node (equal bit for bit).
__global__ void __launch_bounds__(256)
    k_res_restrict2v(Dims df, Dims dc, double h2, const uint8_t* __restrict__ nt,
                     const float* __restrict__ t, const float* __restrict__ z, float* __restrict__ x2,
                     double* __restrict__ sums, float* __restrict__ res, const double* __restrict__ P) {
    double s2 = 0, cnt = 0;
    const unsigned sj = df.nx, sk = (unsigned)df.nx * df.ny;
    const double ih2 = 1.0 / h2;
    long long end;
    for (long long v = xcd_chunk(dc.v, 256, end) + threadIdx.x; v < end; v += 256) {
        int cx, cy, cz;
        vox_xyz(dc, v, cx, cy, cz);
        const unsigned X1 = 2u * (unsigned)cx;
        unsigned Y[4], Z[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            Y[q] = (unsigned)min(max(2 * cy - 1 + q, 0), df.ny - 1) * sj;
            Z[q] = (unsigned)min(max(2 * cz - 1 + q, 0), df.nz - 1) * sk;
        }
        auto pair = [&](const float* a, unsigned o) { return *reinterpret_cast<const float2*>(a + o); };
        // V[c][b][a]: z at fine (X[a], Y[b], Z[c]) with X = clamp(2cx - 1 + a)
        float V[4][4][4];
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const bool full = (b == 1 || b == 2) && (c == 1 || c == 2);
                const bool side = !full && ((b == 1 || b == 2) || (c == 1 || c == 2)) &&
                                  !((b == 0 || b == 3) && (c == 0 || c == 3));
                if (!full && !side) continue;
                const unsigned row = Z[c] + Y[b];
                const float2 m = pair(z, row + X1);
                V[c][b][1] = m.x;
                V[c][b][2] = m.y;
                if (full) {
                    V[c][b][0] = X1 >= 2u ? pair(z, row + X1 - 2u).y : m.x;
                    V[c][b][3] = X1 + 2u < (unsigned)df.nx ? pair(z, row + X1 + 2u).x : m.y;
                }
            }
        double acc = 0;
        int nreg = 0;
#pragma unroll
        for (int c = 1; c <= 2; c++)
#pragma unroll
            for (int b = 1; b <= 2; b++) {
                const unsigned row = Z[c] + Y[b];
                const float2 tv = pair(t, row + X1);
                const unsigned short t2 = *reinterpret_cast<const unsigned short*>(nt + row + X1);
                double2 pv = make_double2(0.0, 0.0);
                if (P) pv = *reinterpret_cast<const double2*>(P + row + X1);
                float ro[2];
#pragma unroll
                for (int a = 1; a <= 2; a++) {
                    auto zat = [&](int aa, int bb, int cc) { return (double)V[cc][bb][aa]; };
                    const bool reg = ((t2 >> (8 * (a - 1))) & 0xFFu) == (unsigned)NT_REG;
                    double az = (zat(a - 1, b, c) + zat(a + 1, b, c) + zat(a, b - 1, c) + zat(a, b + 1, c) +
                                 zat(a, b, c - 1) + zat(a, b, c + 1) - 6.0 * zat(a, b, c)) *
                                ih2;
                    if (P) az -= (a == 1 ? pv.x : pv.y) * zat(a, b, c);
                    const double rv = reg ? (double)(a == 1 ? tv.x : tv.y) - az : 0.0;
                    acc += rv;
                    nreg += reg;
                    ro[a - 1] = (float)rv;
                }
                if (res) *reinterpret_cast<float2*>(res + row + X1) = make_float2(ro[0], ro[1]);
            }
        float xv = nreg ? (float)(-h2 * acc / nreg) : 0.f;
        x2[2 * v] = xv;
        x2[2 * v + 1] = 1.f - (float)nreg / 8.0f;
        if (nreg) {
            s2 += (double)xv * xv;
            cnt += 1;
        }
    }
    block_sum2(s2, cnt);
    if (threadIdx.x == 0) {
        sadd(sums, s2);
        sadd(sums + SSIZE, cnt);
    }
}

// (P y) at node (i, j, k): identity (CF 1) or trilinear from aggregate
// centres (CF 2: per axis weight 3/4 on the own aggregate, 1/4 on the
// neighbour on the node's side, the own one again at the grid's ends)
template <int CF>
__device__ inline double interp(const Dims& dc, const float* __restrict__ y, int i, int j, int k) {
    if (CF == 1) return (double)y[((long long)k * dc.ny + j) * dc.nx + i];
    int c[3] = {i >> 1, j >> 1, k >> 1};
    int o[3] = {(i & 1) ? 1 : -1, (j & 1) ? 1 : -1, (k & 1) ? 1 : -1};
    int n[3] = {dc.nx, dc.ny, dc.nz};
    int cn[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        cn[a] = c[a] + o[a];
        if (cn[a] < 0 || cn[a] >= n[a]) cn[a] = c[a];
    }
    float val = 0.f;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        int xx = (q & 1) ? cn[0] : c[0], yy = (q & 2) ? cn[1] : c[1], zz = (q & 4) ? cn[2] : c[2];
        float w = ((q & 1) ? 0.25f : 0.75f) * ((q & 2) ? 0.25f : 0.75f) * ((q & 4) ? 0.25f : 0.75f);
        val += w * y[((long long)zz * dc.ny + yy) * dc.nx + xx];
    }
    return (double)val;
}

// c = P y on regular nodes (0 elsewhere), and the line-search sums
// out += {c.res, c.A c} (two striped scalars: out, out + SSIZE).  c.A c is taken in edge form,
//   -(1/h^2) [sum over regular-regular edges (c_u - c_v)^2
//             + sum over regular nodes of c_u^2 per non-regular neighbour],
// so each node needs c only at its three forward neighbours (interpolated
// here again rather than read back).
template <int CF>
__global__ void __launch_bounds__(256)
    k_prolong_dots(Dims df, Dims dc, double ih2, const uint8_t* __restrict__ nt,
                   const float* __restrict__ y, const float* __restrict__ res, float* __restrict__ c,
                   double* __restrict__ out, const double* __restrict__ P) {
    const long long sj = df.nx, sk = (long long)df.nx * df.ny;
    double acc[2] = {0, 0};
    long long end;
    for (long long u = xcd_chunk(df.v, 256, end) + threadIdx.x; u < end; u += 256) {
        if (nt[u] != NT_REG) {
            c[u] = 0.f;
            continue;
        }
        int i, j, k;
        vox_xyz(df, u, i, j, k);
        double cu = interp<CF>(dc, y, i, j, k);
        c[u] = (float)cu;
        acc[0] += cu * res[u];
        double e = 0.0;
        const long long st[3] = {1, sj, sk};
        const int di[3] = {1, 0, 0}, dj[3] = {0, 1, 0}, dk[3] = {0, 0, 1};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (nt[u + st[a]] == NT_REG) {
                double cv = interp<CF>(dc, y, i + di[a], j + dj[a], k + dk[a]);
                e += (cu - cv) * (cu - cv);
            } else {
                e += cu * cu;
            }
            if (nt[u - st[a]] != NT_REG) e += cu * cu;
        }
        acc[1] -= e * ih2;
        if (P) acc[1] -= P[u] * cu * cu;  // c.J c = c.A c - sum P c^2
    }
    block_sums<2>(acc);
    if (threadIdx.x == 0) {
        sadd(out, acc[0]);
        sadd(out + SSIZE, acc[1]);
    }
}


// CF 2: c = P y and the line-search sums with one thread per aggregate
// (ceil(n/2) per axis; aggregates past the network grid hold only
// non-regular nodes).  Trilinear interpolation from aggregate centres is
// separable: per axis, fine offset 0 (the aggregate's lower node) takes 3/4
// of the own aggregate and 1/4 of the lower neighbour, offset 1 3/4 own and
// 1/4 upper, offset 2 (the next aggregate's lower node) 3/4 upper and 1/4
// own, neighbours clamped at the grid's ends as in interp.  From the 3x3x3
// aggregate values the thread gets c at fine offsets {0,1,2}^3: its own 8
// nodes and the forward neighbours of the edge form of c.A c, which with c
// zero off the regular nodes is -(1/h^2) sum over grid edges (c_u - c_v)^2.
// EVEN (the fine grid exactly twice the network grid): the node types, res,
// P and c of each fine x-row of the aggregate in 2- and 8-byte pieces; the
// same arithmetic (equal bit for bit).
template <bool EVEN>
__global__ void __launch_bounds__(256)
    k_prolong_dots_agg(Dims df, Dims dc, double ih2, const uint8_t* __restrict__ nt,
                       const float* __restrict__ y, const float* __restrict__ res,
                       float* __restrict__ c, double* __restrict__ out, const double* __restrict__ P) {
    const int ex = (df.nx + 1) >> 1, ey = (df.ny + 1) >> 1, ez = (df.nz + 1) >> 1;
    const long long na = (long long)ex * ey * ez;
    const unsigned sj = df.nx, sk = (unsigned)df.nx * df.ny;
    double acc0 = 0, acc1 = 0, accp = 0;
    long long end;
    for (long long v = xcd_chunk(na, 256, end) + threadIdx.x; v < end; v += 256) {
        const unsigned vv = (unsigned)v;
        const unsigned row = vv / (unsigned)ex;
        const int ax = (int)(vv - row * (unsigned)ex);
        const int az = (int)(row / (unsigned)ey), ay = (int)(row - (unsigned)az * ey);
        const int ox = min(ax, dc.nx - 1), oy = min(ay, dc.ny - 1), oz = min(az, dc.nz - 1);
        const int xs[3] = {max(ox - 1, 0), ox, min(ox + 1, dc.nx - 1)};
        const int ys[3] = {max(oy - 1, 0), oy, min(oy + 1, dc.ny - 1)};
        const int zs[3] = {max(oz - 1, 0), oz, min(oz + 1, dc.nz - 1)};
        // separable passes: [z 0..2][y 0..2][x offset 0..2], then y, then z
        float A[3][3][3];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const float* yr = y + ((long long)zs[k] * dc.ny + ys[j]) * dc.nx;
                float m = yr[xs[0]], o = yr[xs[1]], p = yr[xs[2]];
                A[k][j][0] = 0.75f * o + 0.25f * m;
                A[k][j][1] = 0.75f * o + 0.25f * p;
                A[k][j][2] = 0.75f * p + 0.25f * o;
            }
        float B[3][3][3];  // [z][y offset][x offset]
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                B[k][0][a] = 0.75f * A[k][1][a] + 0.25f * A[k][0][a];
                B[k][1][a] = 0.75f * A[k][1][a] + 0.25f * A[k][2][a];
                B[k][2][a] = 0.75f * A[k][2][a] + 0.25f * A[k][1][a];
            }
        double C[3][3][3];  // [z offset][y offset][x offset], 0 off regular nodes
        const int fx0 = 2 * ax, fy0 = 2 * ay, fz0 = 2 * az;
#pragma unroll
        for (int cc = 0; cc < 3; cc++)
#pragma unroll
            for (int b = 0; b < 3; b++) {
                if (b == 2 && cc == 2) continue;  // not needed
                // EVEN: the row's node types at x offsets 0, 1 (one 2-byte load) and 2
                const int fy = fy0 + b, fz = fz0 + cc;
                const bool row_in = fy < df.ny && fz < df.nz;
                const unsigned rowa = (unsigned)fz * sk + (unsigned)fy * sj + (unsigned)fx0;
                unsigned n01 = 0xFFFFu, n2 = 0xFFu;
                if (EVEN && row_in) {
                    n01 = *reinterpret_cast<const unsigned short*>(nt + rowa);
                    if (b < 2 && cc < 2 && fx0 + 2 < df.nx) n2 = nt[rowa + 2];
                }
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const int nfwd = (a == 2) + (b == 2) + (cc == 2);
                    if (nfwd > 1) continue;  // not needed
                    float w;
                    if (cc == 0) w = 0.75f * B[1][b][a] + 0.25f * B[0][b][a];
                    else if (cc == 1) w = 0.75f * B[1][b][a] + 0.25f * B[2][b][a];
                    else w = 0.75f * B[2][b][a] + 0.25f * B[1][b][a];
                    const int fx = fx0 + a;
                    bool reg = false;
                    if (EVEN) {
                        const unsigned t = a < 2 ? (n01 >> (8 * a)) & 0xFFu : n2;
                        reg = t == (unsigned)NT_REG;
                    } else if (fx < df.nx && row_in) {
                        reg = nt[rowa + (unsigned)a] == NT_REG;
                    }
                    C[cc][b][a] = reg ? (double)w : 0.0;
                }
            }
        double e = 0.0;
#pragma unroll
        for (int cc = 0; cc < 2; cc++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const int fy = fy0 + b, fz = fz0 + cc;
                const unsigned rowa = (unsigned)fz * sk + (unsigned)fy * sj + (unsigned)fx0;
                float2 r2 = make_float2(0.f, 0.f);
                double2 p2 = make_double2(0.0, 0.0);
                if (EVEN) {
                    r2 = *reinterpret_cast<const float2*>(res + rowa);
                    if (P) p2 = *reinterpret_cast<const double2*>(P + rowa);
                }
                float co[2];
#pragma unroll
                for (int a = 0; a < 2; a++) {
                    const int fx = fx0 + a;
                    if (!EVEN && (fx >= df.nx || fy >= df.ny || fz >= df.nz)) continue;
                    const unsigned u = rowa + (unsigned)a;
                    const double cu = C[cc][b][a];
                    co[a] = (float)cu;  // exact: interpolated in FP32
                    if (!EVEN) c[u] = co[a];
                    const double rv = EVEN ? (double)(a == 0 ? r2.x : r2.y) : (double)res[u];
                    acc0 += cu != 0.0 ? cu * rv : 0.0;  // res is defined on regular nodes only
                    if (P && cu != 0.0) accp += (EVEN ? (a == 0 ? p2.x : p2.y) : P[u]) * cu * cu;
                    if (fx + 1 < df.nx) e += (cu - C[cc][b][a + 1]) * (cu - C[cc][b][a + 1]);
                    if (fy + 1 < df.ny) e += (cu - C[cc][b + 1][a]) * (cu - C[cc][b + 1][a]);
                    if (fz + 1 < df.nz) e += (cu - C[cc + 1][b][a]) * (cu - C[cc + 1][b][a]);
                }
                if (EVEN) *reinterpret_cast<float2*>(c + rowa) = make_float2(co[0], co[1]);
            }
        acc1 -= e * ih2;
    }
    acc1 -= accp;  // c.J c = c.A c - sum P c^2 (the Newton Jacobian)
    double acc[2] = {acc0, acc1};
    block_sums<2>(acc);
    if (threadIdx.x == 0) {
        sadd(out, acc[0]);
        sadd(out + SSIZE, acc[1]);
    }
}

}  // namespace

// =========================================================================
// host side: network state
// =========================================================================
constexpr int MAXLV = PIC_UNET_MAX_LEVELS;

struct UNetLayer {
    int cin, cout, taps;  // taps 27 or 1; direct layers: in (cin 2) / out (cout 1)
    size_t woff, boff;    // offsets into the FP32 parameter vector
    bf16* wb = nullptr;   // BF16 weights [cout][taps*cin]
    bf16* wt = nullptr;   // BF16 transposed, tap-flipped
};

struct ULevel {
    Dims d;
    int C;
    bf16 *a = nullptr, *s = nullptr, *p = nullptr;  // p: pooled input (l >= 1)
    bf16 *u = nullptr, *m = nullptr, *h = nullptr;  // decoder (l <= D-2); u on level l+1's grid
    bf16 *ga = nullptr, *gs = nullptr, *gp = nullptr, *gu = nullptr, *gm = nullptr, *gh = nullptr;
};

struct UNet {
    int D = 3, coarsen = 1;
    ULevel lv[MAXLV];
    std::vector<UNetLayer> L;
    int l_in = 0, l_out = 0, l_a[MAXLV] = {}, l_b[MAXLV] = {}, l_u[MAXLV] = {}, l_dec[MAXLV] = {};
    size_t nparam = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    float* x = nullptr;   // network input [V0][2]
    bf16* wi_pad = nullptr;  // input layer weights, BF16 [C0][64] (k = 2 tap + channel)
    float* y = nullptr;   // [V0]
    float* dy = nullptr;  // [V0]
    long long adam_t = 0;
    std::vector<void*> allocs;
    bool trained = false;
};

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

// deterministic initialisation (host, then uploaded): weights and biases
// uniform in +-1/sqrt(fan_in).  (The ReLU-gain bound sqrt(6/fan_in) with zero
// biases trains markedly slower on this regression.)
static void init_params(UNet* u, uint32_t seed, std::vector<float>& h) {
    h.assign(u->nparam, 0.f);
    uint32_t s = seed * 2654435761u + 12345u;
    auto rnd = [&]() {  // xorshift32 -> uniform(-1, 1)
        s ^= s << 13;
        s ^= s >> 17;
        s ^= s << 5;
        return (float)((s >> 8) * (1.0 / 8388608.0) - 1.0);
    };
    for (size_t l = 0; l < u->L.size(); l++) {
        const UNetLayer& ly = u->L[l];
        size_t nw = (size_t)ly.cout * ly.taps * ly.cin;
        float scale = 1.0f / std::sqrt((float)(ly.taps * ly.cin));
        for (size_t q = 0; q < nw; q++) h[ly.woff + q] = scale * rnd();
        for (int q = 0; q < ly.cout; q++) h[ly.boff + q] = scale * rnd();
    }
}

static int unet_build(pic_ctx* ctx, const pic_unet_desc* desc) {
    pic_unet_free(ctx);
    const GridC& g = ctx->g;
    const int D = desc->levels, cf = desc->coarsen;
    if (D < 2 || D > MAXLV) return pic_fail(ctx, PIC_ERR_ARG, "U-Net levels must be 2..5");
    if (cf != 1 && cf != 2) return pic_fail(ctx, PIC_ERR_ARG, "U-Net coarsen must be 1 or 2");
    for (int l = 0; l < D; l++) {
        int c = desc->channels[l];
        if (c < 16 || c % 16 || c > (l == 0 ? 32 : 64))
            return pic_fail(ctx, PIC_ERR_ARG,
                            "U-Net channels must be multiples of 16 (level 0 <= 32, others <= 64)");
    }
    int div = cf << (D - 1);
    if (g.ni % div || g.nj % div || g.nk % div)
        return pic_fail(ctx, PIC_ERR_ARG, "U-Net: node counts must be divisible by coarsen * 2^(levels-1)");
    UNet* u = new UNet();
    ctx->unet = u;
    u->D = D;
    u->coarsen = cf;
    Dims d{g.ni / cf, g.nj / cf, g.nk / cf, 0};
    d.v = (long long)d.nx * d.ny * d.nz;
    for (int l = 0; l < D; l++) {
        u->lv[l].d = l == 0 ? d : scaled_dims(u->lv[l - 1].d, 2);
        u->lv[l].C = desc->channels[l];
    }
    auto add = [&](int cin, int cout, int taps) {
        UNetLayer ly;
        ly.cin = cin;
        ly.cout = cout;
        ly.taps = taps;
        u->L.push_back(ly);
        return (int)u->L.size() - 1;
    };
    const int* C = desc->channels;
    u->l_in = add(2, C[0], 27);
    u->l_b[0] = add(C[0], C[0], 27);
    for (int l = 1; l < D; l++) {
        u->l_a[l] = add(C[l - 1], C[l], 27);
        u->l_b[l] = add(C[l], C[l], 27);
    }
    for (int l = D - 2; l >= 0; l--) {
        u->l_u[l] = add(C[l + 1], C[l], 1);
        u->l_dec[l] = add(C[l], C[l], 27);
    }
    u->l_out = add(C[0], 1, 27);
    size_t off = 0;
    for (auto& ly : u->L) {
        ly.woff = off;
        off += (size_t)ly.cin * ly.cout * ly.taps;
        ly.boff = off;
        off += ly.cout;
    }
    u->nparam = off;
    int rc = 0;
    size_t pb = off * sizeof(float);
    rc |= ualloc(ctx, u, (void**)&u->p, pb);
    rc |= ualloc(ctx, u, (void**)&u->g, pb);
    rc |= ualloc(ctx, u, (void**)&u->m, pb);
    rc |= ualloc(ctx, u, (void**)&u->v, pb);
    for (size_t l = 0; l < u->L.size(); l++) {
        if ((int)l == u->l_in || (int)l == u->l_out) continue;
        size_t nw = (size_t)u->L[l].cin * u->L[l].cout * u->L[l].taps * sizeof(bf16);
        rc |= ualloc(ctx, u, (void**)&u->L[l].wb, nw);
        rc |= ualloc(ctx, u, (void**)&u->L[l].wt, nw);
    }
    rc |= ualloc(ctx, u, (void**)&u->wi_pad, (size_t)C[0] * 64 * sizeof(bf16));
    long long V0 = d.v;
    rc |= ualloc(ctx, u, (void**)&u->x, V0 * 2 * sizeof(float));
    rc |= ualloc(ctx, u, (void**)&u->y, V0 * sizeof(float));
    rc |= ualloc(ctx, u, (void**)&u->dy, V0 * sizeof(float));
    for (int l = 0; l < D && !rc; l++) {
        ULevel& L = u->lv[l];
        long long V = L.d.v, n = V * L.C;
        rc |= ualloc(ctx, u, (void**)&L.a, n * 2);
        rc |= ualloc(ctx, u, (void**)&L.s, n * 2);
        rc |= ualloc(ctx, u, (void**)&L.ga, n * 2);
        if (l >= 1) {
            long long np = V * u->lv[l - 1].C;
            rc |= ualloc(ctx, u, (void**)&L.p, np * 2);
            rc |= ualloc(ctx, u, (void**)&L.gp, np * 2);
        }
        if (l == D - 1) rc |= ualloc(ctx, u, (void**)&L.gs, n * 2);
        if (l <= D - 2) {
            long long nu = u->lv[l + 1].d.v * L.C;
            rc |= ualloc(ctx, u, (void**)&L.u, nu * 2);
            rc |= ualloc(ctx, u, (void**)&L.gu, nu * 2);
            rc |= ualloc(ctx, u, (void**)&L.m, n * 2);
            rc |= ualloc(ctx, u, (void**)&L.h, n * 2);
            rc |= ualloc(ctx, u, (void**)&L.gm, n * 2);
            rc |= ualloc(ctx, u, (void**)&L.gh, n * 2);
        }
    }
    if (rc) {
        pic_unet_free(ctx);
        return PIC_ERR_HIP;
    }
    std::vector<float> h;
    init_params(u, desc->seed, h);
    PIC_HIP(ctx, hipMemcpyAsync(u->p, h.data(), pb, hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP(ctx, hipMemsetAsync(u->m, 0, pb, ctx->stream));
    PIC_HIP(ctx, hipMemsetAsync(u->v, 0, pb, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

// refresh the BF16 weight copies from the FP32 master parameters
static int unet_sync_weights(pic_ctx* ctx, UNet* u) {
    for (size_t l = 0; l < u->L.size(); l++) {
        UNetLayer& ly = u->L[l];
        if (!ly.wb) continue;
        long long n = (long long)ly.cout * ly.taps * ly.cin;
        k_transpose_w<<<grid_for(n, 256), 256, 0, ctx->stream>>>(u->p + ly.woff, ly.cout, ly.cin,
                                                                ly.taps, ly.wb, ly.wt);
    }
    const UNetLayer& li = u->L[u->l_in];
    k_pad_in_w<<<grid_for((size_t)64 * li.cout, 256), 256, 0, ctx->stream>>>(u->p + li.woff, li.cout,
                                                                             u->wi_pad);
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

// returns whether the kernel also wrote the 2x2x2 mean of the output to pool
// (only the register-blocked kernel does; the caller pools otherwise)
static bool conv(pic_ctx* ctx, const Dims& d, const bf16* in, int cin, const bf16* w,
                 const float* bias, bf16* out, int cout, int taps, int relu, const bf16* mask,
                 bf16* pool = nullptr) {
    ConvArgs a{in, cin, w, bias, out, cout, mask, relu, 0, d};
    if (mask) pool = nullptr;  // pooling follows the unmasked forward outputs only
    int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + TY - 1) / TY, ntz = (d.nz + TZ - 1) / TZ;
    int tiles = ntx * nty * ntz;
    // two N-blocks per workgroup when that still leaves >= 512 workgroups
    int nb = (cout % 32 == 0 && tiles * (cout / 32) >= 512) ? 2 : 1;
    dim3 grid(tiles, cout / 16 / nb);
    // register prefetch of the next chunk pays with several chunks despite its
    // registers (measured: 11.5 ms learned solve at 128^3 vs 13.1 ms without)
    const bool pf = cin > 16;
    if (taps == 27) {
        // register-blocked kernel where the shape fits it (every large-grid
        // layer of networks with 32 or 64 channels)
        // 32 output channels: 16x4x4-voxel blocks per wave, two waves along z
        // (measured faster than 16x2x4 blocks with the A-fragment queue)
        if ((cin == 32 || cin == 64) && cout == 32 && d.nx % 16 == 0 && d.ny % 4 == 0 &&
            d.nz % 8 == 0 && (d.nx / 16) * (d.ny / 4) * (d.nz / 8) >= 256) {
            dim3 gr((d.nx / 16) * (d.ny / 4) * (d.nz / 8), 1);
            a.pool = pool;
            if (cin == 32)
                k_conv3_rb<1, 4, 4, 2><<<gr, 256, 0, ctx->stream>>>(a);
            else
                k_conv3_rb<2, 4, 4, 2><<<gr, 256, 0, ctx->stream>>>(a);
            return pool != nullptr;
        }
        // 64 output channels: 16x2x4-voxel blocks (measured faster than 16x2x2
        // with twice the workgroups at 32^3)
        if ((cin == 32 || cin == 64) && cout == 64 && d.nx % 16 == 0 && d.ny % 2 == 0 &&
            d.nz % 4 == 0 && (d.nx / 16) * (d.ny / 2) * (d.nz / 4) >= 256) {
            dim3 gr((d.nx / 16) * (d.ny / 2) * (d.nz / 4), 1);
            a.pool = pool;
            if (cin == 32)
                k_conv3_rb<1, 2, 4, 4><<<gr, 256, 0, ctx->stream>>>(a);
            else
                k_conv3_rb<2, 2, 4, 4><<<gr, 256, 0, ctx->stream>>>(a);
            return pool != nullptr;
        }
        if (tiles * (cout / 16) < 512) {  // small grid: 16x2x2 tiles, taps split over the waves
            int t4 = ntx * ((d.ny + 1) / 2) * ((d.nz + 1) / 2);
            // the pooled output too where a tile is whole pooling cells
            const bool pooling = pool && (cin == 64 || cin == 32) && d.nx % 16 == 0 && d.ny % 2 == 0 &&
                                 d.nz % 2 == 0;
            a.pool = pooling ? pool : nullptr;
            if (cin == 64) {
                k_conv3_small<64><<<dim3(t4, cout / 16), 256, 0, ctx->stream>>>(a);
                return pooling;
            } else if (cin == 32) {
                k_conv3_small<32><<<dim3(t4, cout / 16), 256, 0, ctx->stream>>>(a);
                return pooling;
            }
            else if (pf)
                k_conv3<1, 4, true><<<dim3(t4, cout / 16), 256, 0, ctx->stream>>>(a);
            else
                k_conv3<1, 4, false><<<dim3(t4, cout / 16), 256, 0, ctx->stream>>>(a);
        } else if (nb == 2) {
            dim3 g2(tiles, cout / 32);
            if (pf)
                k_conv3<2, 1, true><<<g2, 256, 0, ctx->stream>>>(a);
            else
                k_conv3<2, 1, false><<<g2, 256, 0, ctx->stream>>>(a);
        } else {
            dim3 g1(tiles, cout / 16);
            if (pf)
                k_conv3<1, 1, true><<<g1, 256, 0, ctx->stream>>>(a);
            else
                k_conv3<1, 1, false><<<g1, 256, 0, ctx->stream>>>(a);
        }
    } else {
        if (nb == 2)
            k_conv_mfma<1, 2><<<grid, 256, 0, ctx->stream>>>(a);
        else
            k_conv_mfma<1, 1><<<grid, 256, 0, ctx->stream>>>(a);
    }
    return false;
}

static int wgrad_tiles(const Dims& d, int* tpb) {
    int ntx = (d.nx + WX - 1) / WX, nty = (d.ny + WY - 1) / WY, ntz = (d.nz + WZ - 1) / WZ;
    int ntiles = ntx * nty * ntz;
    *tpb = std::max(1, ntiles / 512);
    return (ntiles + *tpb - 1) / *tpb;
}

static void wgrad(pic_ctx* ctx, const Dims& d, const bf16* X, int cin, const bf16* dY, int cout,
                  int taps, float* dW, float* db) {
    int tpb;
    int nblk = wgrad_tiles(d, &tpb);
    int combos = taps * (cout / 4) * (cin / 4);
    dim3 grid(nblk, (combos + 1023) / 1024);
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

// forward pass: x (FP32 [V0][2], channel 0 scaled by nrm if given) -> u->y
static int unet_forward(pic_ctx* ctx, UNet* u, const float* x, const double* nrm) {
    const int D = u->D;
    ULevel* lv = u->lv;
    auto& L = u->L;
    float* P = u->p;
    PicProfScope ps(ctx, PIC_K_UNET);
    const Dims& d0 = lv[0].d;
    const int C0 = lv[0].C;
    const UNetLayer& li = L[u->l_in];
    {
        const int tiles = ((d0.nx + TX - 1) / TX) * ((d0.ny + 3) / 4) * ((d0.nz + 3) / 4);
        if (C0 == 32)
            k_conv_in<2><<<tiles, 256, 0, ctx->stream>>>(d0, x, nrm, u->wi_pad, P + li.boff, lv[0].a);
        else
            k_conv_in<1><<<tiles, 256, 0, ctx->stream>>>(d0, x, nrm, u->wi_pad, P + li.boff, lv[0].a);
    }
    // B_l also writes the pooled input of level l + 1 where its kernel can
    bool pooled = conv(ctx, d0, lv[0].a, C0, L[u->l_b[0]].wb, P + L[u->l_b[0]].boff, lv[0].s, C0, 27, 1,
                       nullptr, lv[1].p);
    for (int l = 1; l < D; l++) {
        const Dims& d = lv[l].d;
        int Cp = lv[l - 1].C, Cl = lv[l].C;
        if (!pooled) k_pool<<<ew(d.v * Cp / 8), 256, 0, ctx->stream>>>(d, lv[l - 1].s, lv[l].p, Cp);
        conv(ctx, d, lv[l].p, Cp, L[u->l_a[l]].wb, P + L[u->l_a[l]].boff, lv[l].a, Cl, 27, 1, nullptr);
        pooled = conv(ctx, d, lv[l].a, Cl, L[u->l_b[l]].wb, P + L[u->l_b[l]].boff, lv[l].s, Cl, 27, 1,
                      nullptr, l + 1 < D ? lv[l + 1].p : nullptr);
    }
    for (int l = D - 2; l >= 0; l--) {
        const Dims& d = lv[l].d;
        int Cl = lv[l].C, Cn = lv[l + 1].C;
        const bf16* src = l == D - 2 ? lv[D - 1].s : lv[l + 1].h;
        const UNetLayer& ul = L[u->l_u[l]];
        const Dims& dn = lv[l + 1].d;
        const unsigned nbu = (unsigned)((dn.v + 15) / 16);
        auto up1 = [&](auto ks, auto nb) {
            k_up1_add<decltype(ks)::value, decltype(nb)::value>
                <<<nbu, 256, 0, ctx->stream>>>(dn, src, ul.wb, P + ul.boff, lv[l].s, lv[l].m);
        };
        using I1 = std::integral_constant<int, 1>;
        using I2 = std::integral_constant<int, 2>;
        using I4 = std::integral_constant<int, 4>;
        const bool fused = (Cn == 32 || Cn == 64) && (Cl == 16 || Cl == 32 || Cl == 64);
        if (fused) {
            if (Cn == 32)
                Cl == 16 ? up1(I1{}, I1{}) : Cl == 32 ? up1(I1{}, I2{}) : up1(I1{}, I4{});
            else
                Cl == 16 ? up1(I2{}, I1{}) : Cl == 32 ? up1(I2{}, I2{}) : up1(I2{}, I4{});
        }
        else {  // channel counts of 16 or 48: separate 1x1 convolution and upsampling
            conv(ctx, dn, src, Cn, ul.wb, P + ul.boff, lv[l].u, Cl, 1, 0, nullptr);
            k_up_add<<<ew(d.v * Cl / 8), 256, 0, ctx->stream>>>(d, lv[l].u, lv[l].s, lv[l].m, Cl);
        }
        conv(ctx, d, lv[l].m, Cl, L[u->l_dec[l]].wb, P + L[u->l_dec[l]].boff, lv[l].h, Cl, 27, 1,
             nullptr);
    }
    const UNetLayer& lo = L[u->l_out];
    {
        const int tiles = ((d0.nx + TX - 1) / TX) * ((d0.ny + 3) / 4) * ((d0.nz + 3) / 4);
        if (C0 == 32)
            k_conv_out<32><<<tiles, 256, 0, ctx->stream>>>(d0, lv[0].h, P + lo.woff, P + lo.boff, u->y);
        else
            k_conv_out<16><<<tiles, 256, 0, ctx->stream>>>(d0, lv[0].h, P + lo.woff, P + lo.boff, u->y);
    }
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

template <int C>
static void output_grads(pic_ctx* ctx, UNet* u) {
    const Dims& d0 = u->lv[0].d;
    int tpb;
    int nblk = wgrad_tiles(d0, &tpb);
    const UNetLayer& lo = u->L[u->l_out];
    k_wgrad_direct<1, C, float, bf16><<<nblk, 256, 0, ctx->stream>>>(
        d0, u->lv[0].h, u->dy, u->g + lo.woff, u->g + lo.boff, tpb);
}

template <int C>
static void input_grads(pic_ctx* ctx, UNet* u, const float* x) {
    const Dims& d0 = u->lv[0].d;
    int tpb;
    int nblk = wgrad_tiles(d0, &tpb);
    const UNetLayer& li = u->L[u->l_in];
    k_wgrad_direct<C, 2, bf16, float><<<nblk, 256, 0, ctx->stream>>>(
        d0, x, u->lv[0].ga, u->g + li.woff, u->g + li.boff, tpb);
}

// backward pass from u->dy (FP32 [V0]); accumulates into u->g
static int unet_backward(pic_ctx* ctx, UNet* u, const float* x) {
    const int D = u->D;
    ULevel* lv = u->lv;
    auto& L = u->L;
    float* G = u->g;
    PicProfScope ps(ctx, PIC_K_UNET);
    const Dims& d0 = lv[0].d;
    const int C0 = lv[0].C;
    // output layer
    if (C0 == 16) {
        output_grads<16>(ctx, u);
        k_conv_out_bwd<16><<<ew(d0.v), 256, 0, ctx->stream>>>(d0, u->dy, u->p + L[u->l_out].woff,
                                                              lv[0].h, lv[0].gh);
    } else {
        output_grads<32>(ctx, u);
        k_conv_out_bwd<32><<<ew(d0.v), 256, 0, ctx->stream>>>(d0, u->dy, u->p + L[u->l_out].woff,
                                                              lv[0].h, lv[0].gh);
    }
    // decoder, top down: gh_l is the gradient at h_l's pre-activation
    for (int l = 0; l <= D - 2; l++) {
        const Dims& d = lv[l].d;
        int Cl = lv[l].C, Cn = lv[l + 1].C;
        const UNetLayer& dec = L[u->l_dec[l]];
        wgrad(ctx, d, lv[l].m, Cl, lv[l].gh, Cl, 27, G + dec.woff, G + dec.boff);
        conv(ctx, d, lv[l].gh, Cl, dec.wt, nullptr, lv[l].gm, Cl, 27, 0, nullptr);
        k_up_bwd<<<ew(lv[l + 1].d.v * Cl / 8), 256, 0, ctx->stream>>>(lv[l + 1].d, lv[l].gm, lv[l].gu, Cl);
        const bf16* src = l == D - 2 ? lv[D - 1].s : lv[l + 1].h;
        bf16* gsrc = l == D - 2 ? lv[D - 1].gs : lv[l + 1].gh;
        const UNetLayer& up = L[u->l_u[l]];
        wgrad(ctx, lv[l + 1].d, src, Cn, lv[l].gu, Cl, 1, G + up.woff, G + up.boff);
        conv(ctx, lv[l + 1].d, lv[l].gu, Cl, up.wt, nullptr, gsrc, Cn, 1, 0, src);
    }
    // encoder, bottom up
    for (int l = D - 1; l >= 1; l--) {
        const Dims& d = lv[l].d;
        int Cl = lv[l].C, Cp = lv[l - 1].C;
        bf16* gsl;  // gradient at s_l's pre-activation
        if (l == D - 1) {
            gsl = lv[l].gs;
        } else {
            k_pool_bwd<<<ew(d.v * Cl / 8), 256, 0, ctx->stream>>>(d, lv[l + 1].gp, lv[l].gm, lv[l].s, Cl);
            gsl = lv[l].gm;
        }
        const UNetLayer &B = L[u->l_b[l]], &A = L[u->l_a[l]];
        wgrad(ctx, d, lv[l].a, Cl, gsl, Cl, 27, G + B.woff, G + B.boff);
        conv(ctx, d, gsl, Cl, B.wt, nullptr, lv[l].ga, Cl, 27, 0, lv[l].a);
        wgrad(ctx, d, lv[l].p, Cp, lv[l].ga, Cl, 27, G + A.woff, G + A.boff);
        conv(ctx, d, lv[l].ga, Cl, A.wt, nullptr, lv[l].gp, Cp, 27, 0, nullptr);
    }
    // level 0
    k_pool_bwd<<<ew(d0.v * C0 / 8), 256, 0, ctx->stream>>>(d0, lv[1].gp, lv[0].gm, lv[0].s, C0);
    const UNetLayer& B0 = L[u->l_b[0]];
    wgrad(ctx, d0, lv[0].a, C0, lv[0].gm, C0, 27, G + B0.woff, G + B0.boff);
    conv(ctx, d0, lv[0].gm, C0, B0.wt, nullptr, lv[0].ga, C0, 27, 0, lv[0].a);
    if (C0 == 16)
        input_grads<16>(ctx, u, x);
    else
        input_grads<32>(ctx, u, x);
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

static UNet* get_unet(pic_ctx* ctx) { return ctx ? (UNet*)ctx->unet : nullptr; }

static double count_valid(const float* x2, size_t V) {
    size_t n = 0;
    for (size_t q = 0; q < V; q++) n += x2[2 * q + 1] < 0.999f;
    return (double)std::max<size_t>(1, n);
}

// ---------------------------------------------------------------- ABI
extern "C" int pic_unet_create(pic_ctx* ctx, const pic_unet_desc* desc) {
    if (!ctx || !desc) return PIC_ERR_ARG;
    int rc = unet_build(ctx, desc);
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
    size_t V = (size_t)u->lv[0].d.v;
    PIC_HIP(ctx, hipMemcpyAsync(u->x, in2, 2 * V * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    int rc = unet_forward(ctx, u, u->x, nullptr);
    if (rc) return rc;
    PIC_HIP(ctx, hipMemcpyAsync(out, u->y, V * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PIC_OK;
}

extern "C" int pic_unet_loss_grad(pic_ctx* ctx, const float* in2, const float* target,
                                  double* loss, float* grad) {
    UNet* u = get_unet(ctx);
    if (!u || !in2 || !target) return PIC_ERR_ARG;
    size_t V = (size_t)u->lv[0].d.v;
    float* tdev = nullptr;
    PIC_HIP(ctx, hipMalloc(&tdev, V * sizeof(float)));
    hipError_t e = hipMemcpyAsync(tdev, target, V * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(u->x, in2, 2 * V * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(tdev);
        return pic_hip_check(ctx, e, "U-Net inputs");
    }
    double nvalid = count_valid(in2, V);
    int rc = unet_forward(ctx, u, u->x, nullptr);
    if (!rc) {
        (void)hipMemsetAsync(ctx->scratch, 0, sizeof(double), ctx->stream);
        (void)hipMemsetAsync(u->g, 0, u->nparam * sizeof(float), ctx->stream);
        k_loss<<<ew(V), 256, 0, ctx->stream>>>((long long)V, u->y, tdev, u->x, (float)(1.0 / nvalid),
                                               u->dy, ctx->scratch);
        rc = unet_backward(ctx, u, u->x);
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

// t = a / rms(a) in FP64 (a training residual)
__global__ void __launch_bounds__(256)
    k_norm_t(long long n, const float* __restrict__ a, const double* __restrict__ sumsq, double count,
             double* __restrict__ t) {
    long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    double s = sumsq[0] > 0 ? 1.0 / sqrt(sumsq[0] / count) : 1.0;
    t[e] = (double)a[e] * s;
}

// rho = -eps0 (t - A z) on regular nodes, 0 elsewhere: the plain solver then
// returns e with A e = t - A z (A = 7-point Laplacian / h^2, interior nodes)
__global__ void __launch_bounds__(256)
    k_res_rho(Dims df, double ih2, const uint8_t* __restrict__ nt, const double* __restrict__ t,
              const float* __restrict__ z, double* __restrict__ rho) {
    const long long sj = df.nx, sk = (long long)df.nx * df.ny;
    long long end;
    for (long long u = xcd_chunk(df.v, 256, end) + threadIdx.x; u < end; u += 256) {
        double v = 0.0;
        if (nt[u] == NT_REG) {
            double az = ((double)z[u - 1] + z[u + 1] + z[u - sj] + z[u + sj] + z[u - sk] + z[u + sk] -
                         6.0 * z[u]) *
                        ih2;
            v = -EPS0 * (t[u] - az);
        }
        rho[u] = v;
    }
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

// Coarse correction of the two-level preconditioner: the residual
// res = t - A z (t: the fine right-hand side, FP64; z: the smoothed
// approximation, FP32; res is stored in FP32) is restricted to the network
// grid, the network is applied and its output interpolated to c (FP32).  ws holds four
// striped scalars that the caller has zeroed: the restriction's {x^2, count}
// and then dots = ws + 2 SSIZE, which receives {c.res, c.A c}, from which the
// caller takes the A-norm optimal step omega = (c.res) / (c.A c) for
// z += omega c.  Only for Dirichlet-walled uniform grids (every regular node
// is interior).
int pic_unet_coarse_correct(pic_ctx* ctx, const float* t, const float* z, float* c, float* res,
                            double* ws, const double* P) {
    UNet* u = get_unet(ctx);
    if (!u) return pic_fail(ctx, PIC_ERR_STATE, "no U-Net");
    const GridC& g = ctx->g;
    Dims df{g.ni, g.nj, g.nk, g.nn};
    const Dims& dc = u->lv[0].d;
    double h2 = g.dx[0] * g.dx[0];
    double* sums = ws;  // input normalisation
    double* dots = ws + 2 * SSIZE;
    unsigned nbc = xcd_blocks1(dc.v, 256), nbf = xcd_blocks(df.v, 256);
    if (u->coarsen == 2)
        if (g.ni == 2 * dc.nx && g.nj == 2 * dc.ny && g.nk == 2 * dc.nz)
            k_res_restrict2v<<<nbc, 256, 0, ctx->stream>>>(df, dc, h2, ctx->ntype, t, z, u->x, sums, res, P);
        else
            k_res_restrict<2, float><<<nbc, 256, 0, ctx->stream>>>(df, dc, h2, ctx->ntype, t, z, u->x, sums, res,
                                                                   P);
    else
        k_res_restrict<1, float><<<nbc, 256, 0, ctx->stream>>>(df, dc, h2, ctx->ntype, t, z, u->x, sums, res, P);
    int rc = unet_forward(ctx, u, u->x, sums);
    if (rc) return rc;
    if (u->coarsen == 2) {
        long long na = (long long)((g.ni + 1) / 2) * ((g.nj + 1) / 2) * ((g.nk + 1) / 2);
        if (g.ni == 2 * dc.nx && g.nj == 2 * dc.ny && g.nk == 2 * dc.nz)
            k_prolong_dots_agg<true><<<xcd_blocks1(na, 256), 256, 0, ctx->stream>>>(df, dc, 1.0 / h2, ctx->ntype,
                                                                                    u->y, res, c, dots, P);
        else
            k_prolong_dots_agg<false><<<xcd_blocks1(na, 256), 256, 0, ctx->stream>>>(
                df, dc, 1.0 / h2, ctx->ntype, u->y, res, c, dots, P);
    } else
        k_prolong_dots<1><<<nbf, 256, 0, ctx->stream>>>(df, dc, 1.0 / h2, ctx->ntype, u->y, res, c, dots,
                                                        P);
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

int pic_poisson_internal_sor(pic_ctx* ctx, double tol, int max_it, pic_solver_stats* st);
int pic_learned_presmooth(pic_ctx* ctx, const double* r, float* z);

extern "C" int pic_unet_train(pic_ctx* ctx, const pic_unet_train_desc* td,
                              pic_unet_train_stats* stats) {
    UNet* u = get_unet(ctx);
    if (!u || !td) return pic_fail(ctx, PIC_ERR_STATE, "create the U-Net first");
    if (td->samples < 1 || td->epochs < 1 || td->batch < 1 || td->blobs_max < td->blobs_min ||
        !(td->rel_weight >= 0.0 && td->rel_weight < 1.0))
        return pic_fail(ctx, PIC_ERR_ARG, "bad training parameters");
    const GridC& g = ctx->g;
    if (!uniform_dx(g)) return pic_fail(ctx, PIC_ERR_ARG, "the learned solver needs uniform dx");
    int rc = pic_ensure_ntype(ctx);
    if (rc) return rc;
    long long V = g.nn;
    Dims df{g.ni, g.nj, g.nk, g.nn};
    const Dims& dc = u->lv[0].d;
    long long Vc = dc.v;
    size_t S = (size_t)td->samples;
    float *dens = nullptr, *xs = nullptr, *ts = nullptr;
    double *phi_keep = nullptr, *rho_keep = nullptr, *tf = nullptr;
    float* zp = nullptr;
    Blob* dblobs = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(dens);
        (void)hipFree(xs);
        (void)hipFree(ts);
        (void)hipFree(phi_keep);
        (void)hipFree(rho_keep);
        (void)hipFree(tf);
        (void)hipFree(zp);
        (void)hipFree(dblobs);
    };
    hipError_t e = hipMalloc(&dens, V * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&xs, S * Vc * 2 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&ts, S * Vc * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&phi_keep, V * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&rho_keep, V * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&tf, V * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&zp, V * sizeof(float));
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
    // ---- pairs from the plain solver: a residual t (blobs + white noise), its
    // pre-smoothed remainder res = t - A S(t) restricted to the network grid
    // (the preconditioner's input), and the restricted exact error A^-1 res ----
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
    std::vector<double> nvalid(S, 1.0), tms(S, 0.0);
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
        e = hipMemcpyAsync(dblobs, hb.data(), hb.size() * sizeof(Blob), hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) {
            rc = pic_hip_check(ctx, e, "blobs");
            break;
        }
        float noise = (float)(td->noise * urand());
        k_blobs<<<ew(V), 256, 0, ctx->stream>>>(df, dblobs, nb, noise, td->seed, (uint32_t)s, ctx->ntype,
                                               dens);
        (void)hipMemsetAsync(ss, 0, 3 * sizeof(double), ctx->stream);
        k_sumsq_f<<<std::min<unsigned>(1024, ew(V)), 256, 0, ctx->stream>>>(V, dens, ss);
        k_norm_t<<<ew(V), 256, 0, ctx->stream>>>(V, dens, ss, count, tf);
        rc = pic_learned_presmooth(ctx, tf, zp);
        if (rc) break;
        float* xd = xs + s * Vc * 2;
        float* td_ = ts + s * Vc;
        (void)hipMemsetAsync(ss, 0, 3 * SSIZE * sizeof(double), ctx->stream);  // striped
        unsigned nbc = xcd_blocks1(Vc, 256), nbf = xcd_blocks(V, 256);
        if (u->coarsen == 2)
            k_res_restrict<2><<<nbc, 256, 0, ctx->stream>>>(df, dc, dx2, ctx->ntype, tf, zp, xd, ss, nullptr,
                                                            nullptr);
        else
            k_res_restrict<1><<<nbc, 256, 0, ctx->stream>>>(df, dc, dx2, ctx->ntype, tf, zp, xd, ss, nullptr,
                                                            nullptr);
        k_res_rho<<<nbf, 256, 0, ctx->stream>>>(df, 1.0 / dx2, ctx->ntype, tf, zp, ctx->rho);
        (void)hipMemsetAsync(ctx->phi, 0, V * 8, ctx->stream);
        pic_solver_stats st;
        int r = pic_poisson_internal_sor(ctx, 1e-6, 20000, &st);
        if (r < 0) {
            rc = r;
            break;
        }
        if (u->coarsen == 2)
            k_restrict_target<2><<<nbc, 256, 0, ctx->stream>>>(df, dc, ctx->ntype, ctx->phi, td_, ss);
        else
            k_restrict_target<1><<<nbc, 256, 0, ctx->stream>>>(df, dc, ctx->ntype, ctx->phi, td_, ss);
        k_scale_pair<<<ew(Vc), 256, 0, ctx->stream>>>(Vc, ss, xd, td_);
        std::vector<double> hs(3 * SSIZE);
        (void)hipMemcpyAsync(hs.data(), ss, hs.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            rc = pic_hip_check(ctx, e, "training data");
            break;
        }
        const double x2s = shost(hs.data()), cnt = shost(hs.data() + SSIZE),
                     t2s = shost(hs.data() + 2 * SSIZE);
        nvalid[s] = std::max(1.0, cnt);
        tms[s] = x2s > 0 ? t2s / x2s : 0.0;  // mean square of the scaled target
    }
    // one constant for the whole set brings the targets to unit RMS on average
    // (the solver's line search makes the network's output scale irrelevant).
    // The loss is (nearly) relative per sample: weight 1 / (valid count *
    // (alpha ms + 1 - alpha)), ms = the sample's target mean square over the set
    // mean, so residuals whose error is small -- the rough ones that dominate
    // after the first iterations -- count about as much as smooth ones.
    std::vector<double> inv_w(S, 1.0);
    if (!rc) {
        double m2 = 0;
        for (size_t q = 0; q < S; q++) m2 += tms[q];
        m2 /= (double)S;
        if (m2 > 0)
            k_scale_f<<<ew((long long)S * Vc), 256, 0, ctx->stream>>>((long long)S * Vc,
                                                                     (float)(1.0 / std::sqrt(m2)), ts);
        const double al = td->rel_weight;
        for (size_t q = 0; q < S; q++) {
            double ms = m2 > 0 ? tms[q] / m2 : 1.0;
            inv_w[q] = 1.0 / (nvalid[q] * (al * ms + (1.0 - al)));
        }
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
                const float* xd = xs + s * Vc * 2;
                rc = unet_forward(ctx, u, xd, nullptr);
                if (rc) break;
                k_loss<<<ew(Vc), 256, 0, ctx->stream>>>(Vc, u->y, ts + s * Vc, xd, (float)inv_w[s],
                                                        u->dy, ctx->scratch);
                rc = unet_backward(ctx, u, xd);
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
