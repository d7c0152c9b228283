// pic_solver.hip -- field solve of the gfx950 PIC library.
//
//   k_efield     E = -grad(phi), FP64 math, FP32 float4 output
//                (PotentialSolver.cpp:354-408)
//   k_sor        red-black SOR with the Boltzmann electron term
//                (PotentialSolver.cpp:69-166; ch2/v1/PotentialSolver.cpp:8-87)
//   k_neumann    Neumann face copies (PotentialSolver.cpp:96-107)
//   k_F          Newton residual F = A phi - b - q n0 e^{(phi-phi0)/Te0}/eps0
//                and the Jacobian diagonal shift (PotentialSolver.cpp:189-217)
//   k_pcg1/2     matrix-free Jacobi PCG (PotentialSolver.cpp:243-313), two
//                fused kernels per iteration, scalars kept on the device
//
// The operator rows follow PotentialSolver::buildMatrix
// (PotentialSolver.cpp:414-471): Dirichlet rows are the identity, Neumann face
// rows are one-sided (1/dx)(phi_u - phi_nbr), regular rows are the 7-point
// Laplacian.  Jacobi M = 1/diag(A) of that operator (PotentialSolver.cpp:470).
//
// All stencil kernels walk the grid in contiguous chunks per workgroup and
// map consecutive chunks to the same XCD (workgroups are dealt round-robin
// over the 8 XCDs), so k-neighbour planes are reused from that XCD's L2.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pic_internal.h"

using namespace pic;

namespace {

constexpr int BLOCK = 256;
constexpr int NXCD = 8;

struct Stencil {
    double cx, cy, cz, twos;  // +-1, +-ni, +-ni*nj coefficients; 2(sum 1/dx^2)
    double itwos;             // 1 / twos
    double idx[3];            // 1/dx per axis (Neumann rows)
    long long sj, sk, nn;
    int ni, nj, nk;
};

// gridDim.x is a multiple of NXCD (stencil_blocks), so lb is a bijection
__device__ inline long long chunk_begin(long long nn, long long& end) {
    int nb = gridDim.x;
    int b = blockIdx.x;
    int per_xcd = nb / NXCD;
    int lb = (b % NXCD) * per_xcd + b / NXCD;  // logical, XCD-contiguous
    long long chunk = ((nn + nb - 1) / nb + BLOCK - 1) / BLOCK * BLOCK;
    long long beg = (long long)lb * chunk;
    end = beg + chunk < nn ? beg + chunk : nn;
    return beg;
}

// Node pairs (2h, 2h+1) of every i-line, XCD-contiguous chunks per
// workgroup: f(u, second, red) for the pair's first node u, whether the
// second node exists (odd ni) and whether the first node is red
// ((i + j + k) even; the second node has the other colour).  One (j, k)
// division per pair; a wave's lanes never diverge on the colour.
__device__ inline unsigned pair_count(const Stencil& s) {
    return (unsigned)((s.ni + 1) / 2) * (unsigned)s.nj * (unsigned)s.nk;
}
template <typename F>
__device__ inline void for_pairs(const Stencil& s, F f) {
    const unsigned half = (unsigned)(s.ni + 1) >> 1;
    long long end;
    for (long long m = chunk_begin(pair_count(s), end) + threadIdx.x; m < end; m += BLOCK) {
        unsigned row = (unsigned)m / half;
        unsigned h = (unsigned)m - row * half;
        unsigned k = row / (unsigned)s.nj;
        unsigned j = row - k * (unsigned)s.nj;
        unsigned u = row * (unsigned)s.ni + 2 * h;
        f(u, 2 * h + 1 < (unsigned)s.ni, ((j + k) & 1u) == 0);
    }
}

__device__ inline double block_sum(double v) {
    __shared__ double red[BLOCK / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < BLOCK / 64; w++) t += red[w];
    __syncthreads();
    return t;  // valid in thread 0
}

template <typename T>
__device__ inline T nbr_of(uint8_t t, const T* v, long long u, const Stencil& s) {
    switch (t) {
        case NT_NEU_IM: return v[u + 1];
        case NT_NEU_IP: return v[u - 1];
        case NT_NEU_JM: return v[u + s.sj];
        case NT_NEU_JP: return v[u - s.sj];
        case NT_NEU_KM: return v[u + s.sk];
        default: return v[u - s.sk];
    }
}
__device__ inline double neu_coef(uint8_t t, const Stencil& s) {
    return t <= NT_NEU_IP ? s.idx[0] : (t <= NT_NEU_JP ? s.idx[1] : s.idx[2]);
}
// diag(A) for the Jacobi preconditioner
__device__ inline double diagA(uint8_t t, const Stencil& s) {
    if (t == NT_REG) return -s.twos;
    if (t == NT_DIR) return 1.0;
    return neu_coef(t, s);
}

// (A v)_u for a vector given as a functor of the node index
template <typename T, typename F>
__device__ inline T apply_A(uint8_t t, long long u, const Stencil& s, F v) {
    T vu = v(u);
    if (t == NT_REG)
        return (T)s.cz * (v(u - s.sk) + v(u + s.sk)) + (T)s.cy * (v(u - s.sj) + v(u + s.sj)) +
               (T)s.cx * (v(u - 1) + v(u + 1)) - (T)s.twos * vu;
    if (t == NT_DIR) return vu;
    long long o;
    switch (t) {
        case NT_NEU_IM: o = 1; break;
        case NT_NEU_IP: o = -1; break;
        case NT_NEU_JM: o = s.sj; break;
        case NT_NEU_JP: o = -s.sj; break;
        case NT_NEU_KM: o = s.sk; break;
        default: o = -s.sk; break;
    }
    return (T)neu_coef(t, s) * (vu - v(u + o));
}

// ------------------------------------------------------------------------
// E-field
// ------------------------------------------------------------------------
__device__ inline void efield_node(const GridC& g, const double* __restrict__ phi, float4* __restrict__ ef,
                                   long long u, float (&e)[3]) {
    unsigned uu = (unsigned)u, row = uu / (unsigned)g.ni, kk = row / (unsigned)g.nj;
    int i = (int)(uu - row * (unsigned)g.ni), j = (int)(row - kk * (unsigned)g.nj), k = (int)kk;
    int ijk[3] = {i, j, k};
    int n[3] = {g.ni, g.nj, g.nk};
    long long st[3] = {1, g.ni, (long long)g.ni * g.nj};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double inv2 = 1.0 / (2 * g.dx[d]);
        long long s = st[d];
        double v;
        if (ijk[d] == 0)
            v = (3 * phi[u] - 4 * phi[u + s] + phi[u + 2 * s]) * inv2;
        else if (ijk[d] == n[d] - 1)
            v = (-phi[u - 2 * s] + 4 * phi[u - s] - 3 * phi[u]) * inv2;
        else
            v = (phi[u - s] - phi[u + s]) * inv2;
        e[d] = (float)v;
    }
    ef[u] = make_float4(e[0], e[1], e[2], 0.f);
}

// a few nodes per thread (grid-stride): one maximum reduction per workgroup
__global__ void __launch_bounds__(BLOCK)
    k_efield(GridC g, const double* __restrict__ phi, float4* __restrict__ ef, float* __restrict__ emax) {
    float m[3] = {0.f, 0.f, 0.f};
    for (long long u = (long long)blockIdx.x * BLOCK + threadIdx.x; u < g.nn; u += (long long)gridDim.x * BLOCK) {
        float e[3];
        efield_node(g, phi, ef, u, e);
#pragma unroll
        for (int d = 0; d < 3; d++) m[d] = fmaxf(m[d], fabsf(e[d]));
    }
    block_max3_to(m[0], m[1], m[2], emax);  // the push's bound on the kick
}


// warm start of a time loop's solve: phi = 2 phi - prev, prev = the old phi
__global__ void __launch_bounds__(BLOCK)
    k_extrapolate(long long n, double* __restrict__ phi, double* __restrict__ prev) {
    long long end;
    for (long long m = chunk_begin(n, end) + threadIdx.x; m < end; m += BLOCK) {
        const double p = phi[m];
        phi[m] = 2.0 * p - prev[m];
        prev[m] = p;
    }
}

// ------------------------------------------------------------------------
// red-black SOR
// ------------------------------------------------------------------------
struct SorArgs {
    double omega, ieps, itwos, phi0, n0, inv_te0;
    int color;
};

// The SOR update of a regular node from its neighbour sums along each axis
// (shared by both sweep kernels, rounded the same way: no contraction).
__device__ inline double sor_node(const Stencil& s, const SorArgs& a, double p, double rho_u, double sx,
                                  double sy, double sz) {
#pragma clang fp contract(off)
    double q = rho_u;
    if (a.n0 != 0.0) q -= QE * (a.n0 * exp((p - a.phi0) * a.inv_te0));
    const double nw = (q * a.ieps + sx * s.cx + sy * s.cy + sz * s.cz) * a.itwos;
    return p + a.omega * (nw - p);
}

__global__ void __launch_bounds__(BLOCK)
    k_sor(Stencil s, SorArgs a, const uint8_t* __restrict__ nt, double* __restrict__ phi,
          const double* __restrict__ rho) {
    long long half = (s.ni + 1) / 2;
    long long items = half * s.nj * s.nk;
    long long end;
    for (long long m = chunk_begin(items, end) + threadIdx.x; m < end; m += BLOCK) {
        unsigned row = (unsigned)m / (unsigned)half;
        int h = (int)((unsigned)m - row * (unsigned)half);
        unsigned kk = row / (unsigned)s.nj;
        int j = (int)(row - kk * (unsigned)s.nj), k = (int)kk;
        int i = 2 * h + ((j + k + a.color) & 1);
        if (i >= s.ni) continue;
        long long u = row * s.ni + i;
        if (nt[u] != NT_REG) continue;
        phi[u] = sor_node(s, a, phi[u], rho[u], phi[u - 1] + phi[u + 1], phi[u - s.sj] + phi[u + s.sj],
                          phi[u - s.sk] + phi[u + s.sk]);
    }
}

// Red-black split layout for the SOR sweeps: each colour of a row (j, k)
// stored contiguously, hr = (ni + 1) / 2 entries per row and colour; node
// (i, j, k) of colour c = (i + j + k) & 1 sits at (k nj + j) hr + (i >> 1)
// of colour c's half.  A colour's sweep then reads and writes whole lines
// (the natural layout's stride-2 stores write half of every line twice).
// The neighbours of a node of one colour are the other colour's entries at
// (i -+ 1) >> 1 of its row and at i >> 1 of the rows j -+ 1 and k -+ 1.
__global__ void __launch_bounds__(BLOCK)
    k_split(Stencil s, int hr, long long half, const double* __restrict__ phi, const double* __restrict__ rho,
            const uint8_t* __restrict__ nt, double* __restrict__ sphi, double* __restrict__ srho,
            uint8_t* __restrict__ snt) {
    long long end;
    for (long long u = chunk_begin(s.nn, end) + threadIdx.x; u < end; u += BLOCK) {
        const unsigned row = (unsigned)u / (unsigned)s.ni;
        const int i = (int)((unsigned)u - row * (unsigned)s.ni);
        const unsigned k = row / (unsigned)s.nj, j = row - k * (unsigned)s.nj;
        const long long v = ((i + j + k) & 1) * half + (long long)row * hr + (i >> 1);
        sphi[v] = phi[u];
        srho[v] = rho[u];
        snt[v] = nt[u];
    }
}

__global__ void __launch_bounds__(BLOCK)
    k_unsplit(Stencil s, int hr, long long half, const double* __restrict__ sphi, double* __restrict__ phi) {
    long long end;
    for (long long u = chunk_begin(s.nn, end) + threadIdx.x; u < end; u += BLOCK) {
        const unsigned row = (unsigned)u / (unsigned)s.ni;
        const int i = (int)((unsigned)u - row * (unsigned)s.ni);
        const unsigned k = row / (unsigned)s.nj, j = row - k * (unsigned)s.nj;
        phi[u] = sphi[((i + j + k) & 1) * half + (long long)row * hr + (i >> 1)];
    }
}

// one colour's half sweep on the split layout: the same node update as
// k_sor (sor_node), so the sweeps agree with k_sor bit for bit
__global__ void __launch_bounds__(BLOCK)
    k_sor_split(Stencil s, SorArgs a, int hr, long long half, const uint8_t* __restrict__ snt,
                double* __restrict__ sphi, const double* __restrict__ srho) {
    const int c = a.color;
    double* own = sphi + c * half;
    const double* oth = sphi + (c ^ 1) * half;
    const uint8_t* tn = snt + c * half;
    const double* rh = srho + c * half;
    const long long items = (long long)hr * s.nj * s.nk;
    long long end;
    for (long long m = chunk_begin(items, end) + threadIdx.x; m < end; m += BLOCK) {
        const unsigned row = (unsigned)m / (unsigned)hr;
        const int h = (int)((unsigned)m - row * (unsigned)hr);
        const unsigned k = row / (unsigned)s.nj, j = row - k * (unsigned)s.nj;
        const int i = 2 * h + (int)((j + k + c) & 1);
        if (i >= s.ni || tn[m] != NT_REG) continue;
        const long long rb = (long long)row * hr;
        own[m] = sor_node(s, a, own[m], rh[m], oth[rb + ((i - 1) >> 1)] + oth[rb + ((i + 1) >> 1)],
                          oth[m - hr] + oth[m + hr], oth[m - (long long)s.nj * hr] + oth[m + (long long)s.nj * hr]);
    }
}

// Neumann copies; phase 0 = k faces, 1 = j faces, 2 = i faces, so edge and
// corner chains (an i face reads a j face reads a k face) resolve in order.
__global__ void __launch_bounds__(BLOCK)
    k_neumann(Stencil s, const uint8_t* __restrict__ nt, double* __restrict__ phi, int phase) {
    long long u = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (u >= s.nn) return;
    uint8_t t = nt[u];
    if (t == NT_REG || t == NT_DIR) return;
    int ph = t <= NT_NEU_IP ? 2 : (t <= NT_NEU_JP ? 1 : 0);
    if (ph != phase) return;
    phi[u] = nbr_of<double>(t, phi, u, s);
}

// residual of the (Boltzmann) Poisson problem, three striped scalars at
// out, out + SSIZE, out + 2 SSIZE: sum R^2 over all non-Dirichlet rows
// (reference RMS numerator), sum R^2 over regular rows, sum b^2 over
// regular rows.
__global__ void __launch_bounds__(BLOCK)
    k_resid(Stencil s, SorArgs a, const uint8_t* __restrict__ nt, const double* __restrict__ phi,
            const double* __restrict__ rho, double* __restrict__ out) {
    double r_all = 0, r_reg = 0, b_reg = 0;
    long long end;
    for (long long u = chunk_begin(s.nn, end) + threadIdx.x; u < end; u += BLOCK) {
        uint8_t t = nt[u];
        if (t == NT_DIR) continue;
        double p = phi[u];
        if (t != NT_REG) {
            double R = p - nbr_of<double>(t, phi, u, s);
            r_all += R * R;
            continue;
        }
        double q = rho[u];
        if (a.n0 != 0.0) q -= QE * (a.n0 * exp((p - a.phi0) * a.inv_te0));
        double R = -p * s.twos + q * a.ieps + (phi[u - 1] + phi[u + 1]) * s.cx +
                   (phi[u - s.sj] + phi[u + s.sj]) * s.cy + (phi[u - s.sk] + phi[u + s.sk]) * s.cz;
        r_all += R * R;
        r_reg += R * R;
        double b = rho[u] * a.ieps;
        b_reg += b * b;
    }
    double t0 = block_sum(r_all), t1 = block_sum(r_reg), t2 = block_sum(b_reg);
    if (threadIdx.x == 0) {  // three striped scalars
        sadd(out, t0);
        sadd(out + SSIZE, t1);
        sadd(out + 2 * SSIZE, t2);
    }
}

// ------------------------------------------------------------------------
// Newton residual and Jacobian shift (PotentialSolver.cpp:189-217)
//   b_u = 0 (Neumann), phi_u (Dirichlet), -rho_u/eps0 (regular)
//   F = A phi - b - [regular] q n0 exp((phi-phi0)/Te0)/eps0
//   P = [regular] q n0/(eps0 Te0) exp((phi-phi0)/Te0)
// out[0] += F.F (all rows), out[1] += F.F (regular), out[2] += b.b (regular)
// ------------------------------------------------------------------------
struct NewtonArgs {
    double ieps, phi0, n0, inv_te0;
};

template <typename TF>
__global__ void __launch_bounds__(BLOCK)
    k_F(Stencil s, NewtonArgs a, const uint8_t* __restrict__ nt, const double* __restrict__ phi,
        const double* __restrict__ rho, TF* __restrict__ F, double* __restrict__ P,
        double* __restrict__ out) {
    double f_all = 0, f_reg = 0, b_reg = 0;
    long long end;
    for (long long u = chunk_begin(s.nn, end) + threadIdx.x; u < end; u += BLOCK) {
        uint8_t t = nt[u];
        double Ap = apply_A<double>(t, u, s, [&](long long v) { return phi[v]; });
        double b;
        if (t == NT_REG)
            b = -rho[u] * a.ieps;
        else if (t == NT_DIR)
            b = phi[u];
        else
            b = 0.0;
        double f = Ap - b;
        if (t == NT_REG && a.n0 != 0.0) {
            double ex = exp((phi[u] - a.phi0) * a.inv_te0);
            f -= QE * a.n0 * ex * a.ieps;
            P[u] = QE * a.n0 * a.ieps * a.inv_te0 * ex;
        } else if (P) {
            P[u] = 0.0;
        }
        F[u] = (TF)f;
        f_all += f * f;
        if (t == NT_REG) {
            f_reg += f * f;
            b_reg += b * b;
        }
    }
    double t0 = block_sum(f_all), t1 = block_sum(f_reg), t2 = block_sum(b_reg);
    if (threadIdx.x == 0) {
        atomicAdd(&out[0], t0);
        atomicAdd(&out[1], t1);
        atomicAdd(&out[2], t2);
    }
}

// ------------------------------------------------------------------------
// PCG on J y = F, J = A - diag(P), M = 1/diag(A).  History arrays hold
// dz[t] = d_t.z_t, gs[t] = g_t.s_t, gg[t] = g_t.g_t.
// ------------------------------------------------------------------------
struct PcgHist {
    double* dz;
    double* gs;
    double* gg;
};

template <typename T>
struct PcgVec {
    T *y, *g, *s, *d0, *d1, *z;
    const T* F;
    const double* P;  // may be null
};

// g = -F (y = 0), s = M g, d = -s, gs[0], gg[0]
template <typename T>
__global__ void __launch_bounds__(BLOCK)
    k_pcg0(Stencil s, const uint8_t* __restrict__ nt, PcgVec<T> v, PcgHist h) {
    double gs = 0, gg = 0;
    long long end;
    for (long long u = chunk_begin(s.nn, end) + threadIdx.x; u < end; u += BLOCK) {
        T g = -v.F[u];
        T m = (T)(1.0 / diagA(nt[u], s));
        T sv = m * g;
        v.y[u] = 0;
        v.g[u] = g;
        v.s[u] = sv;
        v.d0[u] = 0;
        gs += (double)g * (double)sv;
        gg += (double)g * (double)g;
    }
    double a = block_sum(gs), b = block_sum(gg);
    if (threadIdx.x == 0) {
        atomicAdd(&h.gs[0], a);
        atomicAdd(&h.gg[0], b);
    }
}

// d_new = beta d_old - s (beta = gs[t]/gs[t-1], 0 at t = 0); z = J d_new,
// by the recurrence J d_new = beta J d_old - J s: the stencil of s alone
// (seven loads instead of fourteen), the previous z in place
template <typename T>
__global__ void __launch_bounds__(BLOCK)
    k_pcg1(Stencil s, const uint8_t* __restrict__ nt, const T* __restrict__ dold,
           T* __restrict__ dnew, const T* __restrict__ sv, T* __restrict__ z,
           const double* __restrict__ P, PcgHist h, int t) {
    double beta = 0.0;
    if (t > 0) {
        double den = h.gs[t - 1];
        beta = den != 0.0 ? h.gs[t] / den : 0.0;
    }
    T bt = (T)beta;
    double acc = 0;
    long long end;
    for (long long u = chunk_begin(s.nn, end) + threadIdx.x; u < end; u += BLOCK) {
        uint8_t ty = nt[u];
        const T su = sv[u];
        T js = apply_A<T>(ty, u, s, [&](long long q) { return sv[q]; });
        if (P) js -= (T)P[u] * su;
        T du = -su, zu = -js;
        if (t > 0) {
            du += bt * dold[u];
            zu += bt * z[u];
        }
        dnew[u] = du;
        z[u] = zu;
        acc += (double)du * (double)zu;
    }
    double a = block_sum(acc);
    if (threadIdx.x == 0) atomicAdd(&h.dz[t], a);
}

// alpha = gs[t]/dz[t]; y += alpha d; g += alpha z; s = M g; gs[t+1], gg[t+1]
template <typename T>
__global__ void __launch_bounds__(BLOCK)
    k_pcg2(Stencil s, const uint8_t* __restrict__ nt, T* __restrict__ y, T* __restrict__ g,
           T* __restrict__ sv, const T* __restrict__ d, const T* __restrict__ z, PcgHist h,
           int t) {
    double den = h.dz[t];
    T alpha = (T)(den != 0.0 ? h.gs[t] / den : 0.0);
    double gs = 0, gg = 0;
    long long end;
    for (long long u = chunk_begin(s.nn, end) + threadIdx.x; u < end; u += BLOCK) {
        T gu = g[u] + alpha * z[u];
        y[u] += alpha * d[u];
        T m = (T)(1.0 / diagA(nt[u], s));
        T su = m * gu;
        g[u] = gu;
        sv[u] = su;
        gs += (double)gu * (double)su;
        gg += (double)gu * (double)gu;
    }
    double a = block_sum(gs), b = block_sum(gg);
    if (threadIdx.x == 0) {
        atomicAdd(&h.gs[t + 1], a);
        atomicAdd(&h.gg[t + 1], b);
    }
}

// phi -= y on non-Dirichlet rows (PotentialSolver.cpp:225-228)
template <typename T>
__global__ void __launch_bounds__(BLOCK)
    k_update_phi(Stencil s, const uint8_t* __restrict__ nt, double* __restrict__ phi,
                 const T* __restrict__ y, double* __restrict__ out) {
    double yy = 0;
    long long end;
    for (long long u = chunk_begin(s.nn, end) + threadIdx.x; u < end; u += BLOCK) {
        if (nt[u] == NT_DIR) continue;
        double yv = (double)y[u];
        phi[u] -= yv;
        yy += yv * yv;
    }
    double a = block_sum(yy);
    if (threadIdx.x == 0) atomicAdd(out, a);
}

Stencil make_stencil(const GridC& g, bool axis_swap) {
    Stencil s;
    double d2x = 1.0 / (g.dx[0] * g.dx[0]), d2y = 1.0 / (g.dx[1] * g.dx[1]),
           d2z = 1.0 / (g.dx[2] * g.dx[2]);
    s.cx = axis_swap ? d2z : d2x;
    s.cy = d2y;
    s.cz = axis_swap ? d2x : d2z;
    s.twos = 2.0 * (d2x + d2y + d2z);
    s.itwos = 1.0 / s.twos;
    for (int d = 0; d < 3; d++) s.idx[d] = 1.0 / g.dx[d];
    s.sj = g.ni;
    s.sk = (long long)g.ni * g.nj;
    s.nn = g.nn;
    s.ni = g.ni;
    s.nj = g.nj;
    s.nk = g.nk;
    return s;
}

// kernels ending in a single-address grid-wide sum: fewer, longer workgroups
// (eight items per thread) keep the per-workgroup atomics few
unsigned stencil_blocks_red(long long items) {
    long long b = (items + BLOCK * 8 - 1) / (BLOCK * 8);
    if (b > 2048) b = 2048;
    b = (b + NXCD - 1) / NXCD * NXCD;
    return (unsigned)b;
}

unsigned stencil_blocks(long long items) {
    // about two items per thread: enough workgroups in flight to cover HBM latency
    long long b = (items + BLOCK * 2 - 1) / (BLOCK * 2);
    if (b > 16384) b = 16384;
    b = (b + NXCD - 1) / NXCD * NXCD;  // multiple of NXCD, at least NXCD
    return (unsigned)b;
}

// workgroups of the node-pair kernels (for_pairs): one pair per thread
// (measured faster than two)
unsigned pair_blocks(const Stencil& s) { return stencil_blocks(s.nn); }

}  // namespace

// ------------------------------------------------------------------------
extern "C" int pic_efield(pic_ctx* ctx) {
    if (!ctx) return PIC_ERR_ARG;
    {
        PicProfScope ps(ctx, PIC_K_EFIELD);
        PIC_HIP(ctx, hipMemsetAsync(ctx->emax, 0, 4 * sizeof(float), ctx->stream));
        k_efield<<<std::min(grid_for((size_t)ctx->g.nn, BLOCK), 2048u), BLOCK, 0, ctx->stream>>>(
            ctx->g, ctx->phi, ctx->ef, ctx->emax);
    }
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

// any Neumann face node (a face node that is neither an object node nor
// flagged Dirichlet); the node types are current
static bool has_neumann(pic_ctx* ctx) {
    if (pic_ensure_ntype(ctx) != PIC_OK) return true;
    return ctx->has_neumann;
}

static int residual_norms(pic_ctx* ctx, const Stencil& st, const SorArgs& a, double out[3],
                          const double* phi = nullptr) {
    double* sums = ctx->scratch + LEARNED_WS + 5 * SSIZE;  // three striped scalars
    PIC_HIP(ctx, hipMemsetAsync(sums, 0, 3 * SSIZE * sizeof(double), ctx->stream));
    k_resid<<<stencil_blocks(st.nn), BLOCK, 0, ctx->stream>>>(st, a, ctx->ntype, phi ? phi : ctx->phi,
                                                             ctx->rho, sums);
    PIC_HIP(ctx, hipGetLastError());
    std::vector<double> h(3 * SSIZE);
    PIC_HIP(ctx, hipMemcpyAsync(h.data(), sums, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    PIC_HIP(ctx, pic_spin_wait(ctx->stream));
    for (int q = 0; q < 3; q++) out[q] = shost(h.data() + q * SSIZE);
    return PIC_OK;
}

static double crit_value(const pic_solver_params* p, const double r[3], double nn) {
    if (p->tol_kind == PIC_TOL_REF_RMS) return std::sqrt(r[0] / nn);
    return r[2] > 0 ? std::sqrt(r[1] / r[2]) : std::sqrt(r[1]);
}

static int solve_sor(pic_ctx* ctx, const pic_solver_params* p, pic_solver_stats* st) {
    const GridC& g = ctx->g;
    Stencil s = make_stencil(g, false);
    SorArgs a;
    a.omega = p->omega > 0 ? p->omega : 1.4;
    a.ieps = 1.0 / EPS0;
    a.itwos = 1.0 / s.twos;
    a.phi0 = p->phi0;
    a.n0 = p->n0;
    a.inv_te0 = 1.0 / (p->Te0 != 0 ? p->Te0 : 1.0);
    bool neu = has_neumann(ctx);
    int check = p->check_every > 0 ? p->check_every : 25;
    long long half_items = (long long)((g.ni + 1) / 2) * g.nj * g.nk;
    // without Neumann faces the sweeps run on the red-black split layout
    // (the same values as the natural layout's launches)
    const bool split = !neu && !(p->flags & PIC_SOLVE_SOR_PLAIN);
    const int hr = (g.ni + 1) / 2;
    const long long half = (long long)hr * g.nj * g.nk;
    double *sphi = nullptr, *srho = nullptr;
    uint8_t* snt = nullptr;
    if (split) {
        const size_t bytes = (size_t)half * 2 * (2 * sizeof(double) + 1);
        if (ctx->sor_split_bytes < bytes) {
            PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
            (void)hipFree(ctx->sor_split);
            ctx->sor_split = nullptr;
            ctx->sor_split_bytes = 0;
            PIC_HIP(ctx, hipMalloc(&ctx->sor_split, bytes));
            ctx->sor_split_bytes = bytes;
        }
        sphi = reinterpret_cast<double*>(ctx->sor_split);
        srho = sphi + 2 * half;
        snt = reinterpret_cast<uint8_t*>(srho + 2 * half);
        k_split<<<stencil_blocks(s.nn), BLOCK, 0, ctx->stream>>>(s, hr, half, ctx->phi, ctx->rho, ctx->ntype, sphi,
                                                                 srho, snt);
    }
    double r[3] = {0, 0, 0};
    int it;
    bool conv = false, fresh = false;  // fresh: r holds the residual of the current potential
    double crit = 0;
    for (it = 0; it < p->max_it; it++) {
        {
            PicProfScope ps(ctx, PIC_K_SOR);
            for (int c = 0; c < 2; c++) {
                a.color = c;
                if (split)
                    k_sor_split<<<stencil_blocks(half), BLOCK, 0, ctx->stream>>>(s, a, hr, half, snt, sphi, srho);
                else
                    k_sor<<<stencil_blocks(half_items), BLOCK, 0, ctx->stream>>>(s, a, ctx->ntype, ctx->phi,
                                                                                ctx->rho);
            }
            if (neu)
                for (int ph = 0; ph < 3; ph++)
                    k_neumann<<<grid_for((size_t)g.nn, BLOCK), BLOCK, 0, ctx->stream>>>(s, ctx->ntype, ctx->phi,
                                                                                        ph);
        }
        PIC_HIP(ctx, hipGetLastError());
        fresh = false;
        if (it % check == 0) {  // PotentialSolver.cpp:124
            if (split) k_unsplit<<<stencil_blocks(s.nn), BLOCK, 0, ctx->stream>>>(s, hr, half, sphi, ctx->phi);
            int rc = residual_norms(ctx, s, a, r);
            if (rc) return rc;
            fresh = true;
            crit = crit_value(p, r, (double)g.nn);
            if (crit < p->tol) {
                conv = true;
                break;
            }
        }
    }
    if (split && !fresh) k_unsplit<<<stencil_blocks(s.nn), BLOCK, 0, ctx->stream>>>(s, hr, half, sphi, ctx->phi);
    if (!fresh) {
        int rc = residual_norms(ctx, s, a, r);
        if (rc) return rc;
    }
    st->converged = conv;
    st->iterations = conv ? it + 1 : it;
    st->newton_iterations = 0;
    st->ref_l2 = std::sqrt(r[0] / (double)g.nn);
    st->rel_residual = r[2] > 0 ? std::sqrt(r[1] / r[2]) : std::sqrt(r[1]);
    return conv ? PIC_OK : PIC_NOT_CONVERGED;
}

// inner PCG on J y = F (vectors of type T); stops when ||g|| <= target or at
// max_it; returns iterations used.
template <typename T>
static int pcg_inner(pic_ctx* ctx, const Stencil& s, T* base, const double* P, double target2,
                     int max_it, int check, int* iters, double* gg_last) {
    size_t nn = (size_t)s.nn;
    PcgVec<T> v;
    v.F = base;
    v.y = base + nn;
    v.g = base + 2 * nn;
    v.s = base + 3 * nn;
    v.d0 = base + 4 * nn;
    v.d1 = base + 5 * nn;
    v.z = base + 6 * nn;
    v.P = P;
    double* hist = (double*)(base + 7 * nn);
    hist = (double*)(((uintptr_t)hist + 255) & ~(uintptr_t)255);
    PcgHist h;
    h.dz = hist;
    h.gs = hist + (max_it + 2);
    h.gg = hist + 2 * (max_it + 2);
    PIC_HIP(ctx, hipMemsetAsync(hist, 0, 3 * (size_t)(max_it + 2) * sizeof(double), ctx->stream));
    unsigned nb = stencil_blocks_red(s.nn);
    {
        // set-up launches are not counted as PCG iterations
        k_pcg0<T><<<nb, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, v, h);
    }
    PIC_HIP(ctx, hipGetLastError());
    std::vector<double> hg(max_it + 2);
    PIC_HIP(ctx, hipMemcpyAsync(hg.data(), h.gg, sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    PIC_HIP(ctx, pic_spin_wait(ctx->stream));
    *gg_last = hg[0];
    if (hg[0] <= target2) {  // already converged (zero RHS included)
        *iters = 0;
        return PIC_OK;
    }
    T* dold = v.d0;
    T* dnew = v.d1;
    int t = 0, last_checked = 0;
    while (t < max_it) {
        int upto = std::min(max_it, t + check);
        for (; t < upto; t++) {  // one profiled "launch" = one PCG iteration (2 kernels)
            PicProfScope ps(ctx, PIC_K_PCG);
            k_pcg1<T><<<nb, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, dold, dnew, v.s, v.z, P, h,
                                                     t);
            k_pcg2<T><<<nb, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, v.y, v.g, v.s, dnew, v.z, h,
                                                     t);
            std::swap(dold, dnew);
        }
        PIC_HIP(ctx, hipGetLastError());
        PIC_HIP(ctx, hipMemcpyAsync(hg.data() + last_checked + 1, h.gg + last_checked + 1,
                                    (size_t)(t - last_checked) * sizeof(double),
                                    hipMemcpyDeviceToHost, ctx->stream));
        PIC_HIP(ctx, pic_spin_wait(ctx->stream));
        for (int q = last_checked + 1; q <= t; q++) {
            if (!(hg[q] == hg[q])) return pic_fail(ctx, PIC_ERR_NUMERIC, "PCG produced NaN");
            if (hg[q] <= target2) {
                *iters = t;
                *gg_last = hg[q];
                return PIC_OK;
            }
        }
        last_checked = t;
        *gg_last = hg[t];
    }
    *iters = t;
    return PIC_NOT_CONVERGED;
}

// Newton + PCG (PotentialSolver.cpp:178-241).  The linear case (n0 == 0)
// needs a single Newton step; each Newton step restarts PCG from y = 0.
static int solve_pcg(pic_ctx* ctx, const pic_solver_params* p, pic_solver_stats* st, bool mixed) {
    const GridC& g = ctx->g;
    size_t nn = (size_t)g.nn;
    Stencil s = make_stencil(g, (p->flags & PIC_SOLVE_AXIS_SWAP) != 0);
    bool nonlinear = p->n0 != 0.0;
    int max_it = p->max_it > 0 ? p->max_it : 10000;
    int check = p->check_every > 0 ? p->check_every : 8;
    size_t tsz = mixed ? sizeof(float) : sizeof(double);
    size_t bytes = 7 * nn * tsz + 256 + 3 * (size_t)(max_it + 2) * sizeof(double);
    size_t pbytes = nonlinear ? nn * sizeof(double) : 0;
    int rc = pic_ensure_work(ctx, bytes + pbytes + 256);
    if (rc) return rc;
    char* wbase = (char*)ctx->work;
    double* P = nonlinear ? (double*)(wbase + ((bytes + 255) & ~(size_t)255)) : nullptr;
    NewtonArgs na;
    na.ieps = 1.0 / EPS0;
    na.phi0 = p->phi0;
    na.n0 = p->n0;
    na.inv_te0 = 1.0 / (p->Te0 != 0 ? p->Te0 : 1.0);
    unsigned nb = stencil_blocks_red(s.nn);
    int total_it = 0, newton = 0;
    bool conv = false;
    double fr[3];
    const int newton_max = nonlinear ? 20 : 1000;  // PotentialSolver.cpp:179
    for (newton = 0; newton < newton_max; newton++) {
        PIC_HIP(ctx, hipMemsetAsync(ctx->scratch, 0, 4 * sizeof(double), ctx->stream));
        {
            // set-up launches are not counted as PCG iterations
            if (mixed)
                k_F<float><<<nb, BLOCK, 0, ctx->stream>>>(s, na, ctx->ntype, ctx->phi, ctx->rho,
                                                          (float*)wbase, P, ctx->scratch);
            else
                k_F<double><<<nb, BLOCK, 0, ctx->stream>>>(s, na, ctx->ntype, ctx->phi, ctx->rho,
                                                           (double*)wbase, P, ctx->scratch);
        }
        PIC_HIP(ctx, hipGetLastError());
        PIC_HIP(ctx, hipMemcpyAsync(fr, ctx->scratch, sizeof(fr), hipMemcpyDeviceToHost,
                                    ctx->stream));
        PIC_HIP(ctx, pic_spin_wait(ctx->stream));
        // target on the inner residual g (g = -F at y = 0)
        double target2;
        if (p->tol_kind == PIC_TOL_REF_RMS)
            target2 = p->tol * p->tol * (double)nn;
        else
            target2 = p->tol * p->tol * (fr[2] > 0 ? fr[2] : 1.0);
        if (!nonlinear && fr[0] <= target2) {
            conv = true;
            break;
        }
        // relative criterion on the nonlinear problem: the FP64 residual F
        // of the current phi over the regular rows (the Newton iterate is
        // judged by the residual, not by the size of its last step)
        if (nonlinear && p->tol_kind == PIC_TOL_REL && fr[1] <= target2) {
            conv = true;
            break;
        }
        double inner_target2 = target2;
        if (mixed) {  // FP32 inner solve: reduce the residual by ~1e-3 per pass
            double floor2 = fr[0] * 1e-6;
            inner_target2 = std::max(target2 * 0.25, floor2);
        } else if (nonlinear && p->tol_kind == PIC_TOL_REL) {
            // inexact Newton: each linear solve to 1e-4 of the current
            // residual, or below the final tolerance
            inner_target2 = std::max(target2 * 0.25, fr[0] * 1e-8);
        }
        int its = 0;
        double gg = 0;
        if (mixed)
            rc = pcg_inner<float>(ctx, s, (float*)wbase, P, inner_target2, max_it - total_it, check,
                                  &its, &gg);
        else
            rc = pcg_inner<double>(ctx, s, (double*)wbase, P, inner_target2, max_it - total_it,
                                   check, &its, &gg);
        total_it += its;
        if (rc < 0) return rc;
        PIC_HIP(ctx, hipMemsetAsync(ctx->scratch, 0, sizeof(double), ctx->stream));
        {
            // set-up launches are not counted as PCG iterations
            if (mixed)
                k_update_phi<float><<<nb, BLOCK, 0, ctx->stream>>>(
                    s, ctx->ntype, ctx->phi, (const float*)wbase + nn, ctx->scratch);
            else
                k_update_phi<double><<<nb, BLOCK, 0, ctx->stream>>>(
                    s, ctx->ntype, ctx->phi, (const double*)wbase + nn, ctx->scratch);
        }
        PIC_HIP(ctx, hipGetLastError());
        double yy;
        PIC_HIP(ctx, hipMemcpyAsync(&yy, ctx->scratch, sizeof(double), hipMemcpyDeviceToHost,
                                    ctx->stream));
        PIC_HIP(ctx, pic_spin_wait(ctx->stream));
        if (nonlinear && p->tol_kind == PIC_TOL_REF_RMS) {
            if (std::sqrt(yy / (double)nn) < 1e-3) {  // NR_TOL, PotentialSolver.cpp:180, :229-233
                conv = true;
                newton++;
                break;
            }
        }
        if (total_it >= max_it) break;
    }
    SorArgs a;
    a.omega = 1;
    a.ieps = 1.0 / EPS0;
    a.itwos = 1.0 / s.twos;
    a.phi0 = p->phi0;
    a.n0 = p->n0;
    a.inv_te0 = na.inv_te0;
    double r[3];
    Stencil sphys = make_stencil(g, (p->flags & PIC_SOLVE_AXIS_SWAP) != 0);
    rc = residual_norms(ctx, sphys, a, r);
    if (rc) return rc;
    st->converged = conv;
    st->iterations = total_it;
    st->newton_iterations = newton;
    st->ref_l2 = std::sqrt(r[0] / (double)nn);
    st->rel_residual = r[2] > 0 ? std::sqrt(r[1] / r[2]) : std::sqrt(r[1]);
    return conv ? PIC_OK : PIC_NOT_CONVERGED;
}

// ------------------------------------------------------------------------
// Flexible CG with the learned two-level preconditioner.  FCG(1) (Notay
// 2000): p_k = z_k - (z_k.A p_{k-1}) / (p_{k-1}.A p_{k-1}) p_{k-1}, which keeps
// CG's descent when the preconditioner varies from step to step.  On a cold
// start (phi = 0) the first step's correction is the preconditioned residual
// -- the network's prediction -- scaled by the optimal step length, so the
// network supplies the initial guess.  Scalars stay on the device in
// per-iteration history slots; the host reads the residual history every
// few iterations.  Plain PCG is the fallback and the final residual is
// recomputed in FP64.
//
// Preconditioner M(r): LEARNED_PRE_SWEEPS red-black Gauss-Seidel sweeps on
// A z = r from z = 0 (red, black), the network's coarse correction of the
// remaining residual, LEARNED_POST_SWEEPS sweeps in the reverse colour order
// (black, red).  A sweep costs ~17 us at 128^3, an iteration ~290 us: three
// sweeps on each side take the 128^3 time loop's solves from 5 iterations
// (one sweep each side) to 2 and cold solves from 11 to 6.  The sweeps after
// the first and the output z are stored in FP32 (FP64 arithmetic).
// ------------------------------------------------------------------------
bool pic_unet_ready(pic_ctx* ctx);
int pic_unet_coarse_correct(pic_ctx* ctx, const float* t, const float* z, float* c, float* res,
                            double* ws, const double* P);

namespace {

constexpr int LEARNED_PRE_SWEEPS = 3, LEARNED_POST_SWEEPS = 3;

// per-iteration striped scalars (slot k at base + k * SSIZE)
struct FcgHist {
    double* rz;  // r_k . z_k
    double* pq;  // p_k . A p_k
    double* zq;  // z_{k+1} . A p_k
    double* rr;  // r_{k+1} . r_{k+1}
};

// The neighbour offsets of node u clamped only at the array's ends: the full
// offsets on every regular (interior) node, in bounds on every node, and
// independent of the node type, so a kernel's neighbour loads are issued
// together with its node-type load instead of after it (values on
// non-regular nodes are discarded)
struct NbrOff {
    unsigned mi, pi, mj, pj, mk, pk;
};
__device__ inline NbrOff nbr_off(const Stencil& s, unsigned u) {
    const unsigned n = (unsigned)s.nn, sj = (unsigned)s.sj, sk = (unsigned)s.sk;
    NbrOff o;
    o.mi = u >= 1u ? 1u : 0u;
    o.pi = u + 1u < n ? 1u : 0u;
    o.mj = u >= sj ? sj : 0u;
    o.pj = u + sj < n ? sj : 0u;
    o.mk = u >= sk ? sk : 0u;
    o.pk = u + sk < n ? sk : 0u;
    return o;
}
// 7-point neighbour sum of f at node u, weighted per axis, with nbr_off's
// offsets (the full stencil on regular nodes)
template <typename F>
__device__ inline double nbr_sum_any(const Stencil& s, unsigned u, F f) {
    const NbrOff o = nbr_off(s, u);
    return (f(u - o.mi) + f(u + o.pi)) * s.cx + (f(u - o.mj) + f(u + o.pj)) * s.cy +
           (f(u - o.mk) + f(u + o.pk)) * s.cz;
}

// r = b - A phi on regular rows (0 elsewhere); out += {r.r, b.b, #(phi != 0)}
__global__ void __launch_bounds__(BLOCK)
    k_fcg_resid(Stencil s, const uint8_t* __restrict__ nt, const double* __restrict__ phi,
                const double* __restrict__ rho, double ieps, double* __restrict__ r,
                double* __restrict__ out) {
    double rr = 0, bb = 0, cold = 0;
    auto ph = [&](unsigned w) { return phi[w]; };
    long long end;
    for (long long m = chunk_begin(s.nn, end) + threadIdx.x; m < end; m += BLOCK) {
        const unsigned u = (unsigned)m;
        const bool reg = nt[u] == NT_REG;  // regular nodes are interior (Dirichlet walls)
        const double pu = phi[u];
        const double b = -rho[u] * ieps;
        const double rv = reg ? b - (nbr_sum_any(s, u, ph) - s.twos * pu) : 0.0;
        if (reg) {
            bb += b * b;
            cold += pu != 0.0;
        }
        r[u] = rv;
        rr += rv * rv;
    }
    double a = block_sum(rr), b2 = block_sum(bb), c = block_sum(cold);
    if (threadIdx.x == 0) {  // striped: out, out + SSIZE, out + 2 SSIZE
        sadd(out, a);
        sadd(out + SSIZE, b2);
        sadd(out + 2 * SSIZE, c);
    }
}

// p_k = z_k + beta p_{k-1} (beta = -zq[k-1] / pq[k-1]; p_0 = z_0), q = A p_k on
// regular rows (by the recurrence A p_k = A z_k + beta A p_{k-1}); pq[k] +=
// p_k.q.  z and p vanish on non-regular nodes, and
// regular nodes are interior (Dirichlet walls), so q is the 7-point stencil.
// Neighbour offsets are zeroed on non-regular nodes so that every load is in
// bounds and the loop has no branches.
__global__ void __launch_bounds__(BLOCK)
    k_fcg_pq(Stencil s, const uint8_t* __restrict__ nt, const float* __restrict__ z,
             const double* __restrict__ pold, double* __restrict__ pnew, double* __restrict__ q,
             FcgHist h, int k, const double* __restrict__ P) {
    double beta = 0.0;
    if (k > 0) {
        double den = sget(h.pq + (k - 1) * SSIZE);
        beta = den != 0.0 ? -sget(h.zq + (k - 1) * SSIZE) / den : 0.0;
    }
    const bool first = k == 0;
    double acc = 0;
    long long end;
    for (long long m = chunk_begin(s.nn, end) + threadIdx.x; m < end; m += BLOCK) {
        const unsigned u = (unsigned)m;
        const bool reg = nt[u] == NT_REG;
        const NbrOff o = nbr_off(s, u);
        // q = A p = A z + beta A p_old: the stencil of z alone (seven loads
        // instead of fourteen), the previous q in place
        const double zu = z[u];
        double az = ((double)z[u - o.mk] + z[u + o.pk]) * s.cz + ((double)z[u - o.mj] + z[u + o.pj]) * s.cy +
                    ((double)z[u - o.mi] + z[u + o.pi]) * s.cx - s.twos * zu;
        if (P) az -= P[u] * zu;  // J = A - diag(P) (the Newton Jacobian)
        double pu = zu, qv = az;
        if (!first) {
            pu += beta * pold[u];
            qv += beta * q[u];
        }
        qv = reg ? qv : 0.0;
        pnew[u] = pu;
        q[u] = qv;
        acc += pu * qv;
    }
    double a = block_sum(acc);
    if (threadIdx.x == 0) sadd(h.pq + k * SSIZE, a);
}

// alpha = rz[k] / pq[k]; x += alpha p; r -= alpha q; rr[k] += r.r
__global__ void __launch_bounds__(BLOCK)
    k_fcg_update(long long n, FcgHist h, int k, double* __restrict__ x, const double* __restrict__ p,
                 double* __restrict__ r, const double* __restrict__ q) {
    double den = sget(h.pq + k * SSIZE);
    double alpha = den != 0.0 ? sget(h.rz + k * SSIZE) / den : 0.0;
    double rr = 0;
    long long end;
    for (long long m = chunk_begin(n, end) + threadIdx.x; m < end; m += BLOCK) {
        const unsigned u = (unsigned)m;
        x[u] += alpha * p[u];
        double rv = r[u] - alpha * q[u];
        r[u] = rv;
        rr += rv * rv;
    }
    double a = block_sum(rr);
    if (threadIdx.x == 0) sadd(h.rr + k * SSIZE, a);
}


// One red-black Gauss-Seidel sweep on A z = r from z = 0, in one pass: the
// red half-sweep gives z = -r / (2 sum 1/dx^2) on red regular nodes
// ((i + j + k) even); the black half-sweep that follows reads only those red
// values, so it is written directly in r.  Non-regular nodes get 0.  z is
// FP32 (the preconditioner's vectors are stored in FP32, its arithmetic is
// FP64); r32 (optional) receives r in FP32 for the closing sweeps.
// 1 / the diagonal of the operator's regular row: 2 sum 1/dx^2, plus P_u
// for the Newton Jacobian J = A - diag(P)
__device__ inline double inv_diag(const Stencil& s, const double* __restrict__ P, unsigned u) {
    return P ? 1.0 / (s.twos + P[u]) : s.itwos;
}

__global__ void __launch_bounds__(BLOCK)
    k_presmooth(Stencil s, const uint8_t* __restrict__ nt, const double* __restrict__ r,
                float* __restrict__ z, double* __restrict__ clear, int nclear,
                const double* __restrict__ P, float* __restrict__ r32) {
    // the coarse correction's striped sums start from zero (no memset node)
    for (int q = blockIdx.x * BLOCK + threadIdx.x; q < nclear; q += gridDim.x * BLOCK) clear[q] = 0.0;
    // r / diag on regular nodes, 0 elsewhere (a red value is -that)
    auto rr = [&](unsigned v) {
        double rv = r[v] * inv_diag(s, P, v);
        return nt[v] == NT_REG ? rv : 0.0;
    };
    auto node = [&](unsigned u, bool red) {
        const bool reg = nt[u] == NT_REG;
        const double id = inv_diag(s, P, u);
        double zu;
        if (red)
            zu = -r[u] * id;
        else
            zu = (-nbr_sum_any(s, u, rr) - r[u]) * id;
        z[u] = reg ? (float)zu : 0.f;
        if (r32) r32[u] = (float)r[u];
    };
    for_pairs(s, [&](unsigned u, bool second, bool red) {
        node(u, red);
        if (second) node(u + 1, !red);
    });
}

// The closing red half-sweep of the preconditioner fused with FCG's dot
// products: red regular nodes get the Gauss-Seidel update (their neighbours
// are black and final); then zq += z.q (when q is given) and rz += z.r.
// z vanishes on non-regular nodes.  z and r in FP32 (r32: r rounded), the
// sums in FP64 of the stored values.
__global__ void __launch_bounds__(BLOCK)
    k_gs_red_dots(Stencil s, const uint8_t* __restrict__ nt, const float* __restrict__ r,
                  float* __restrict__ z, const double* __restrict__ q, double* __restrict__ zq,
                  double* __restrict__ rz, const double* __restrict__ P) {
    double a = 0, b = 0;
    auto zv = [&](unsigned w) { return (double)z[w]; };
    auto node = [&](unsigned u, bool red) {
        double zu = z[u];
        if (red) {
            const bool reg = nt[u] == NT_REG;
            double upd = (nbr_sum_any(s, u, zv) - (double)r[u]) * inv_diag(s, P, u);
            if (reg) {
                const float zf = (float)upd;
                z[u] = zf;
                zu = zf;
            }
        }
        if (q) a += zu * q[u];
        b += zu * (double)r[u];
    };
    for_pairs(s, [&](unsigned u, bool second, bool red) {
        node(u, red);
        if (second) node(u + 1, !red);
    });
    double ta = block_sum(a), tb = block_sum(b);
    if (threadIdx.x == 0) {
        if (zq) sadd(zq, ta);
        sadd(rz, tb);
    }
}

// the coarse correction fused with the next (black) half-sweep:
// z = zin + omega c on red nodes, the Gauss-Seidel update from those values on
// black regular nodes, 0 on non-regular nodes; omega = d[0] / d[1] (A-norm
// optimal step along c).  zin and c (FP32) vanish on non-regular nodes; r
// and z in FP32.
__global__ void __launch_bounds__(BLOCK)
    k_gs_corr(Stencil s, const uint8_t* __restrict__ nt, const float* __restrict__ r,
              const float* __restrict__ zin, const float* __restrict__ c,
              const double* __restrict__ d, float* __restrict__ z, const double* __restrict__ P) {
    const double d1 = sget(d + SSIZE);
    const double om = d1 != 0.0 ? sget(d) / d1 : 0.0;
    auto zc = [&](unsigned w) { return (double)zin[w] + om * c[w]; };
    auto node = [&](unsigned u, bool red) {
        const bool reg = nt[u] == NT_REG;
        double zu;
        if (red)
            zu = zc(u);
        else
            zu = (nbr_sum_any(s, u, zc) - (double)r[u]) * inv_diag(s, P, u);
        z[u] = reg ? (float)zu : 0.f;
    };
    for_pairs(s, [&](unsigned u, bool second, bool red) {
        node(u, red);
        if (second) node(u + 1, !red);
    });
}

// one Gauss-Seidel half-sweep in place: the regular nodes of one colour
// (red: (i + j + k) even) get (sum of neighbours - r) / diag; their
// neighbours have the other colour, so the update is race-free.  r in FP32
// (the solver's rounded copy) or FP64.
template <typename RT>
__global__ void __launch_bounds__(BLOCK)
    k_gs_half(Stencil s, const uint8_t* __restrict__ nt, const RT* __restrict__ r,
              float* __restrict__ z, int red_colour, const double* __restrict__ P) {
    auto zv = [&](unsigned w) { return (double)z[w]; };
    auto node = [&](unsigned u, bool red) {
        if (red != (red_colour != 0)) return;
        const double upd = (nbr_sum_any(s, u, zv) - (double)r[u]) * inv_diag(s, P, u);
        if (nt[u] == NT_REG) z[u] = (float)upd;
    };
    for_pairs(s, [&](unsigned u, bool second, bool red) {
        node(u, red);
        if (second) node(u + 1, !red);
    });
}

// k_gs_half<float> on four consecutive nodes of an i-line per thread (ni a
// multiple of 4): the z rows around the group come in 16-byte loads (the
// i-neighbours at the group's ends as single words), r and the node types
// likewise, and the group goes back in one 16-byte store with the other
// colour's two values unchanged.  The same arithmetic as k_gs_half (equal
// bit for bit); every value an update reads has the other colour, which no
// thread writes.
__global__ void __launch_bounds__(BLOCK)
    k_gs_half4(Stencil s, const uint8_t* __restrict__ nt, const float* __restrict__ r,
               float* __restrict__ z, int red_colour, const double* __restrict__ P) {
    const long long ng = s.nn / 4;
    const unsigned n = (unsigned)s.nn, sj = (unsigned)s.sj, sk = (unsigned)s.sk;
    const unsigned want = red_colour ? 0u : 1u;  // parity (i + j + k) & 1 of the updated nodes
    long long end;
    for (long long m = chunk_begin(ng, end) + threadIdx.x; m < end; m += BLOCK) {
        const unsigned u0 = 4u * (unsigned)m;
        const unsigned row = u0 / (unsigned)s.ni;
        const unsigned i0 = u0 - row * (unsigned)s.ni;
        const unsigned k = row / (unsigned)s.nj, j = row - k * (unsigned)s.nj;
        const unsigned p0 = (i0 + j + k) & 1u;
        // in-bounds addresses everywhere; only regular (interior) nodes are
        // updated, and their neighbours are the unclamped ones
        const float4 zc = *reinterpret_cast<const float4*>(z + u0);
        const float zl = z[u0 >= 1u ? u0 - 1u : u0];
        const float zr = z[u0 + 4u < n ? u0 + 4u : u0];
        const float4 ym = *reinterpret_cast<const float4*>(z + (u0 >= sj ? u0 - sj : u0));
        const float4 yp = *reinterpret_cast<const float4*>(z + (u0 + sj < n ? u0 + sj : u0));
        const float4 zm = *reinterpret_cast<const float4*>(z + (u0 >= sk ? u0 - sk : u0));
        const float4 zp = *reinterpret_cast<const float4*>(z + (u0 + sk < n ? u0 + sk : u0));
        const float4 r4 = *reinterpret_cast<const float4*>(r + u0);
        const uint32_t t4 = *reinterpret_cast<const uint32_t*>(nt + u0);
        const float c[4] = {zc.x, zc.y, zc.z, zc.w};
        const float xm[4] = {zl, zc.x, zc.y, zc.z}, xp[4] = {zc.y, zc.z, zc.w, zr};
        const float ymv[4] = {ym.x, ym.y, ym.z, ym.w}, ypv[4] = {yp.x, yp.y, yp.z, yp.w};
        const float zmv[4] = {zm.x, zm.y, zm.z, zm.w}, zpv[4] = {zp.x, zp.y, zp.z, zp.w};
        const float rv[4] = {r4.x, r4.y, r4.z, r4.w};
        float out[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            out[a] = c[a];
            if (((p0 + (unsigned)a) & 1u) != want || ((t4 >> (8 * a)) & 0xFFu) != (uint32_t)NT_REG) continue;
            const double sum = ((double)xm[a] + (double)xp[a]) * s.cx + ((double)ymv[a] + (double)ypv[a]) * s.cy +
                               ((double)zmv[a] + (double)zpv[a]) * s.cz;
            out[a] = (float)((sum - (double)rv[a]) * inv_diag(s, P, u0 + (unsigned)a));
        }
        *reinterpret_cast<float4*>(z + u0) = make_float4(out[0], out[1], out[2], out[3]);
    }
}

// The four-node forms of k_gs_corr and k_gs_red_dots (ni a multiple of 4):
// a group of four consecutive nodes of an i-line per thread, its
// neighbour rows in 16-byte loads, one 16-byte store; the same arithmetic
// (z equal bit for bit; the dot products sum in another order).
struct Group4 {
    unsigned u0, p0;  // first node, its parity (i + j + k) & 1
    unsigned xm, xp, ym, yp, zm, zp;  // clamped addresses: u0 - 1, u0 + 4, u0 -+ sj, u0 -+ sk
};
__device__ inline Group4 group4(const Stencil& s, long long m) {
    Group4 g;
    const unsigned n = (unsigned)s.nn, sj = (unsigned)s.sj, sk = (unsigned)s.sk;
    g.u0 = 4u * (unsigned)m;
    const unsigned row = g.u0 / (unsigned)s.ni;
    const unsigned i0 = g.u0 - row * (unsigned)s.ni;
    const unsigned k = row / (unsigned)s.nj, j = row - k * (unsigned)s.nj;
    g.p0 = (i0 + j + k) & 1u;
    g.xm = g.u0 >= 1u ? g.u0 - 1u : g.u0;
    g.xp = g.u0 + 4u < n ? g.u0 + 4u : g.u0;
    g.ym = g.u0 >= sj ? g.u0 - sj : g.u0;
    g.yp = g.u0 + sj < n ? g.u0 + sj : g.u0;
    g.zm = g.u0 >= sk ? g.u0 - sk : g.u0;
    g.zp = g.u0 + sk < n ? g.u0 + sk : g.u0;
    return g;
}
// a vector's values around a group: [a] = at u0 + a, xm/xp the i-neighbours
// of node a, ym/yp/zm/zp the rows on either side
struct Around4 {
    float c[4], xm[4], xp[4], ym[4], yp[4], zm[4], zp[4];
};
__device__ inline void load_around4(const float* __restrict__ v, const Group4& g, Around4& a) {
    const float4 c = *reinterpret_cast<const float4*>(v + g.u0);
    const float l = v[g.xm], r = v[g.xp];
    const float4 ym = *reinterpret_cast<const float4*>(v + g.ym);
    const float4 yp = *reinterpret_cast<const float4*>(v + g.yp);
    const float4 zm = *reinterpret_cast<const float4*>(v + g.zm);
    const float4 zp = *reinterpret_cast<const float4*>(v + g.zp);
    const float cc[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        a.c[q] = cc[q];
        a.xm[q] = q == 0 ? l : cc[q - 1];
        a.xp[q] = q == 3 ? r : cc[q + 1];
    }
    a.ym[0] = ym.x; a.ym[1] = ym.y; a.ym[2] = ym.z; a.ym[3] = ym.w;
    a.yp[0] = yp.x; a.yp[1] = yp.y; a.yp[2] = yp.z; a.yp[3] = yp.w;
    a.zm[0] = zm.x; a.zm[1] = zm.y; a.zm[2] = zm.z; a.zm[3] = zm.w;
    a.zp[0] = zp.x; a.zp[1] = zp.y; a.zp[2] = zp.z; a.zp[3] = zp.w;
}
unsigned group4_blocks(const Stencil& s) {
    long long b = (s.nn / 4 + BLOCK - 1) / BLOCK;
    return (unsigned)((b + NXCD - 1) / NXCD * NXCD);
}

__global__ void __launch_bounds__(BLOCK)
    k_gs_corr4(Stencil s, const uint8_t* __restrict__ nt, const float* __restrict__ r,
               const float* __restrict__ zin, const float* __restrict__ c,
               const double* __restrict__ d, float* __restrict__ z, const double* __restrict__ P) {
    const double d1 = sget(d + SSIZE);
    const double om = d1 != 0.0 ? sget(d) / d1 : 0.0;
    long long end;
    for (long long m = chunk_begin(s.nn / 4, end) + threadIdx.x; m < end; m += BLOCK) {
        const Group4 g = group4(s, m);
        Around4 zi, cc;
        load_around4(zin, g, zi);
        load_around4(c, g, cc);
        const float4 r4 = *reinterpret_cast<const float4*>(r + g.u0);
        const uint32_t t4 = *reinterpret_cast<const uint32_t*>(nt + g.u0);
        const float rv[4] = {r4.x, r4.y, r4.z, r4.w};
        auto zc = [&](float a, float b) { return (double)a + om * b; };
        float out[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const bool red = ((g.p0 + (unsigned)a) & 1u) == 0u;
            const bool reg = ((t4 >> (8 * a)) & 0xFFu) == (uint32_t)NT_REG;
            double zu;
            if (red)
                zu = zc(zi.c[a], cc.c[a]);
            else
                zu = ((zc(zi.xm[a], cc.xm[a]) + zc(zi.xp[a], cc.xp[a])) * s.cx +
                      (zc(zi.ym[a], cc.ym[a]) + zc(zi.yp[a], cc.yp[a])) * s.cy +
                      (zc(zi.zm[a], cc.zm[a]) + zc(zi.zp[a], cc.zp[a])) * s.cz - (double)rv[a]) *
                     inv_diag(s, P, g.u0 + (unsigned)a);
            out[a] = reg ? (float)zu : 0.f;
        }
        *reinterpret_cast<float4*>(z + g.u0) = make_float4(out[0], out[1], out[2], out[3]);
    }
}

__global__ void __launch_bounds__(BLOCK)
    k_gs_red_dots4(Stencil s, const uint8_t* __restrict__ nt, const float* __restrict__ r,
                   float* __restrict__ z, const double* __restrict__ q, double* __restrict__ zq,
                   double* __restrict__ rz, const double* __restrict__ P) {
    double acc_a = 0, acc_b = 0;
    long long end;
    for (long long m = chunk_begin(s.nn / 4, end) + threadIdx.x; m < end; m += BLOCK) {
        const Group4 g = group4(s, m);
        Around4 zz;
        load_around4(z, g, zz);
        const float4 r4 = *reinterpret_cast<const float4*>(r + g.u0);
        const uint32_t t4 = *reinterpret_cast<const uint32_t*>(nt + g.u0);
        const float rv[4] = {r4.x, r4.y, r4.z, r4.w};
        double qv[4] = {0, 0, 0, 0};
        if (q) {
            const double2 q0 = *reinterpret_cast<const double2*>(q + g.u0);
            const double2 q1 = *reinterpret_cast<const double2*>(q + g.u0 + 2);
            qv[0] = q0.x; qv[1] = q0.y; qv[2] = q1.x; qv[3] = q1.y;
        }
        float out[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            out[a] = zz.c[a];
            const bool red = ((g.p0 + (unsigned)a) & 1u) == 0u;
            const bool reg = ((t4 >> (8 * a)) & 0xFFu) == (uint32_t)NT_REG;
            if (red && reg) {
                const double sum = ((double)zz.xm[a] + (double)zz.xp[a]) * s.cx +
                                   ((double)zz.ym[a] + (double)zz.yp[a]) * s.cy +
                                   ((double)zz.zm[a] + (double)zz.zp[a]) * s.cz;
                out[a] = (float)((sum - (double)rv[a]) * inv_diag(s, P, g.u0 + (unsigned)a));
            }
            const double zu = out[a];
            if (q) acc_a += zu * qv[a];
            acc_b += zu * (double)rv[a];
        }
        *reinterpret_cast<float4*>(z + g.u0) = make_float4(out[0], out[1], out[2], out[3]);
    }
    double ta = block_sum(acc_a), tb = block_sum(acc_b);
    if (threadIdx.x == 0) {
        if (zq) sadd(zq, ta);
        sadd(rz, tb);
    }
}

// The four-node form of k_presmooth without the Jacobian's diagonal (the
// plain operator): r and the node types of the group's neighbour rows in
// 16-byte loads, z and r32 of the group in 16-byte stores; the same
// arithmetic (equal bit for bit).
__device__ inline void load_around4d(const double* __restrict__ v, const Group4& g, double (&c)[4],
                                     double (&xm)[4], double (&xp)[4], double (&ym)[4], double (&yp)[4],
                                     double (&zm)[4], double (&zp)[4]) {
    auto row = [&](unsigned a, double (&o)[4]) {
        const double2 h0 = *reinterpret_cast<const double2*>(v + a);
        const double2 h1 = *reinterpret_cast<const double2*>(v + a + 2);
        o[0] = h0.x; o[1] = h0.y; o[2] = h1.x; o[3] = h1.y;
    };
    row(g.u0, c);
    row(g.ym, ym);
    row(g.yp, yp);
    row(g.zm, zm);
    row(g.zp, zp);
    const double l = v[g.xm], r = v[g.xp];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        xm[q] = q == 0 ? l : c[q - 1];
        xp[q] = q == 3 ? r : c[q + 1];
    }
}
__device__ inline void load_nt4(const uint8_t* __restrict__ nt, const Group4& g, uint32_t (&c)[4],
                                uint32_t (&xm)[4], uint32_t (&xp)[4], uint32_t (&ym)[4], uint32_t (&yp)[4],
                                uint32_t (&zm)[4], uint32_t (&zp)[4]) {
    auto row = [&](unsigned a, uint32_t (&o)[4]) {
        const uint32_t w = *reinterpret_cast<const uint32_t*>(nt + a);
#pragma unroll
        for (int q = 0; q < 4; q++) o[q] = (w >> (8 * q)) & 0xFFu;
    };
    row(g.u0, c);
    row(g.ym, ym);
    row(g.yp, yp);
    row(g.zm, zm);
    row(g.zp, zp);
    const uint32_t l = nt[g.xm], r = nt[g.xp];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        xm[q] = q == 0 ? l : c[q - 1];
        xp[q] = q == 3 ? r : c[q + 1];
    }
}

__global__ void __launch_bounds__(BLOCK)
    k_presmooth4(Stencil s, const uint8_t* __restrict__ nt, const double* __restrict__ r,
                 float* __restrict__ z, double* __restrict__ clear, int nclear, float* __restrict__ r32) {
    for (int q = blockIdx.x * BLOCK + threadIdx.x; q < nclear; q += gridDim.x * BLOCK) clear[q] = 0.0;
    long long end;
    for (long long m = chunk_begin(s.nn / 4, end) + threadIdx.x; m < end; m += BLOCK) {
        const Group4 g = group4(s, m);
        double rc[4], rxm[4], rxp[4], rym[4], ryp[4], rzm[4], rzp[4];
        uint32_t tc[4], txm[4], txp[4], tym[4], typ[4], tzm[4], tzp[4];
        load_around4d(r, g, rc, rxm, rxp, rym, ryp, rzm, rzp);
        load_nt4(nt, g, tc, txm, txp, tym, typ, tzm, tzp);
        // r / diag on regular nodes, 0 elsewhere
        auto rr = [&](double v, uint32_t t) {
            const double rv = v * s.itwos;
            return t == (uint32_t)NT_REG ? rv : 0.0;
        };
        float out[4], o32[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const bool red = ((g.p0 + (unsigned)a) & 1u) == 0u;
            double zu;
            if (red)
                zu = -rc[a] * s.itwos;
            else
                zu = (-((rr(rxm[a], txm[a]) + rr(rxp[a], txp[a])) * s.cx +
                        (rr(rym[a], tym[a]) + rr(ryp[a], typ[a])) * s.cy +
                        (rr(rzm[a], tzm[a]) + rr(rzp[a], tzp[a])) * s.cz) -
                      rc[a]) *
                     s.itwos;
            out[a] = tc[a] == (uint32_t)NT_REG ? (float)zu : 0.f;
            o32[a] = (float)rc[a];
        }
        *reinterpret_cast<float4*>(z + g.u0) = make_float4(out[0], out[1], out[2], out[3]);
        *reinterpret_cast<float4*>(r32 + g.u0) = make_float4(o32[0], o32[1], o32[2], o32[3]);
    }
}

// The four-node form of k_fcg_pq (ni a multiple of 4): z's neighbour rows,
// p_old, q and P in 16-byte loads, p_new and q in 16-byte stores; the same
// arithmetic per node (p and q equal bit for bit; pq sums in another order).
__global__ void __launch_bounds__(BLOCK)
    k_fcg_pq4(Stencil s, const uint8_t* __restrict__ nt, const float* __restrict__ z,
              const double* __restrict__ pold, double* __restrict__ pnew, double* __restrict__ q,
              FcgHist h, int k, const double* __restrict__ P) {
    double beta = 0.0;
    if (k > 0) {
        double den = sget(h.pq + (k - 1) * SSIZE);
        beta = den != 0.0 ? -sget(h.zq + (k - 1) * SSIZE) / den : 0.0;
    }
    const bool first = k == 0;
    double acc = 0;
    auto ld4 = [&](const double* v, unsigned a, double (&o)[4]) {
        const double2 h0 = *reinterpret_cast<const double2*>(v + a);
        const double2 h1 = *reinterpret_cast<const double2*>(v + a + 2);
        o[0] = h0.x; o[1] = h0.y; o[2] = h1.x; o[3] = h1.y;
    };
    long long end;
    for (long long m = chunk_begin(s.nn / 4, end) + threadIdx.x; m < end; m += BLOCK) {
        const Group4 g = group4(s, m);
        Around4 zz;
        load_around4(z, g, zz);
        const uint32_t t4 = *reinterpret_cast<const uint32_t*>(nt + g.u0);
        double po[4] = {0, 0, 0, 0}, qo[4] = {0, 0, 0, 0}, pv[4] = {0, 0, 0, 0};
        if (!first) {
            ld4(pold, g.u0, po);
            ld4(q, g.u0, qo);
        }
        if (P) ld4(P, g.u0, pv);
        double pn[4], qn[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const bool reg = ((t4 >> (8 * a)) & 0xFFu) == (uint32_t)NT_REG;
            const double zu = zz.c[a];
            double az = ((double)zz.zm[a] + zz.zp[a]) * s.cz + ((double)zz.ym[a] + zz.yp[a]) * s.cy +
                        ((double)zz.xm[a] + zz.xp[a]) * s.cx - s.twos * zu;
            if (P) az -= pv[a] * zu;
            double pu = zu, qv = az;
            if (!first) {
                pu += beta * po[a];
                qv += beta * qo[a];
            }
            qv = reg ? qv : 0.0;
            pn[a] = pu;
            qn[a] = qv;
            acc += pu * qv;
        }
        *reinterpret_cast<double2*>(pnew + g.u0) = make_double2(pn[0], pn[1]);
        *reinterpret_cast<double2*>(pnew + g.u0 + 2) = make_double2(pn[2], pn[3]);
        *reinterpret_cast<double2*>(q + g.u0) = make_double2(qn[0], qn[1]);
        *reinterpret_cast<double2*>(q + g.u0 + 2) = make_double2(qn[2], qn[3]);
    }
    double a = block_sum(acc);
    if (threadIdx.x == 0) sadd(h.pq + k * SSIZE, a);
}

// b = a in FP64
__global__ void __launch_bounds__(BLOCK) k_widen(long long n, const float* __restrict__ a, double* __restrict__ b) {
    for (long long i = (long long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (long long)gridDim.x * BLOCK)
        b[i] = (double)a[i];
}

}  // namespace

// one half-sweep of the learned cycle on FP32 vectors: the four-node kernel
// where i-lines are a multiple of 4 long, else the per-pair one
static void gs_half32(pic_ctx* ctx, const Stencil& s, const float* r32, float* z, int red_colour,
                      const double* P) {
    if (s.ni % 4 == 0) {
        k_gs_half4<<<group4_blocks(s), BLOCK, 0, ctx->stream>>>(s, ctx->ntype, r32, z, red_colour, P);
    } else {
        k_gs_half<float><<<pair_blocks(s), BLOCK, 0, ctx->stream>>>(s, ctx->ntype, r32, z, red_colour, P);
    }
}

// z = LEARNED_PRE_SWEEPS red-black Gauss-Seidel sweeps on A z = r from z = 0
// (device vectors): the learned preconditioner's pre-smoother, also used for
// its training pairs
int pic_learned_presmooth(pic_ctx* ctx, const double* r, float* z) {
    Stencil s = make_stencil(ctx->g, false);
    unsigned nb = pair_blocks(s);
    k_presmooth<<<nb, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, r, z, nullptr, 0, nullptr, nullptr);
    for (int e = 1; e < LEARNED_PRE_SWEEPS; e++) {
        k_gs_half<double><<<nb, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, r, z, 1, nullptr);
        k_gs_half<double><<<nb, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, r, z, 0, nullptr);
    }
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

// z = M(r) in FP32; tmp holds the four FP32 vectors zpre, c, res and r32
// (r rounded, read by the closing sweeps).  The closing half-sweep also
// accumulates *zq += z.q (q may be null) and *rz += z.r; both must be zeroed
// by the caller.
static int learned_precond(pic_ctx* ctx, const Stencil& s, const double* rin, float* zout,
                           float* tmp, const double* q, double* zq, double* rz,
                           const double* P = nullptr) {
    unsigned nb = pair_blocks(s);
    float* zpre = tmp;
    float* c = tmp + s.nn;
    float* res = c + s.nn;
    float* r32 = res + s.nn;
    double* ws = ctx->scratch + LEARNED_WS;  // the coarse correction's striped sums
    double* dots = ws + 2 * SSIZE;
    const bool v4 = s.ni % 4 == 0;  // the four-node kernels
    if (v4 && !P)
        k_presmooth4<<<group4_blocks(s), BLOCK, 0, ctx->stream>>>(s, ctx->ntype, rin, zpre, ws, 4 * SSIZE, r32);
    else
        k_presmooth<<<nb, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, rin, zpre, ws, 4 * SSIZE, P, r32);
    for (int e = 1; e < LEARNED_PRE_SWEEPS; e++) {
        gs_half32(ctx, s, r32, zpre, 1, P);
        gs_half32(ctx, s, r32, zpre, 0, P);
    }
    int e = pic_unet_coarse_correct(ctx, r32, zpre, c, res, ws, P);
    if (e) return e;
    // the closing sweeps: the first black half in k_gs_corr, the last red
    // half in k_gs_red_dots, the half-sweeps between them in place
    if (v4)
        k_gs_corr4<<<group4_blocks(s), BLOCK, 0, ctx->stream>>>(s, ctx->ntype, r32, zpre, c, dots, zout, P);
    else
        k_gs_corr<<<nb, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, r32, zpre, c, dots, zout, P);
    for (int e = 1; e < LEARNED_POST_SWEEPS; e++) {
        gs_half32(ctx, s, r32, zout, 1, P);
        gs_half32(ctx, s, r32, zout, 0, P);
    }
    if (v4)
        k_gs_red_dots4<<<group4_blocks(s), BLOCK, 0, ctx->stream>>>(s, ctx->ntype, r32, zout, q, zq, rz, P);
    else
        k_gs_red_dots<<<nb, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, r32, zout, q, zq, rz, P);
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

static int learned_check(pic_ctx* ctx) {
    if (!pic_unet_ready(ctx))
        return pic_fail(ctx, PIC_ERR_STATE, "learned solver: no trained network for this grid");
    if (has_neumann(ctx))
        return pic_fail(ctx, PIC_ERR_ARG, "learned solver: Dirichlet-walled problems only");
    return PIC_OK;
}

// z = M(r) for host vectors (node order), the preconditioner alone
extern "C" int pic_learned_precondition(pic_ctx* ctx, const double* r, double* z) {
    if (!ctx || !r || !z) return PIC_ERR_ARG;
    int rc = pic_ensure_ntype(ctx);
    if (rc) return rc;
    rc = learned_check(ctx);
    if (rc) return rc;
    size_t nn = (size_t)ctx->g.nn;
    rc = pic_ensure_work(ctx, 5 * nn * sizeof(double));
    if (rc) return rc;
    double* rd = ctx->work;
    float* z32 = reinterpret_cast<float*>(rd + nn);
    double* rz = ctx->scratch + LEARNED_WS + 4 * SSIZE;  // striped
    PIC_HIP(ctx, hipMemcpyAsync(rd, r, nn * 8, hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP(ctx, hipMemsetAsync(rz, 0, SSIZE * sizeof(double), ctx->stream));
    rc = learned_precond(ctx, make_stencil(ctx->g, false), rd, z32, reinterpret_cast<float*>(rd + 2 * nn),
                         nullptr, nullptr, rz);
    if (rc) return rc;
    k_widen<<<grid_for(nn, BLOCK), BLOCK, 0, ctx->stream>>>((long long)nn, z32, rd);  // r is spent
    PIC_HIP(ctx, hipGetLastError());
    PIC_HIP(ctx, hipMemcpyAsync(z, rd, nn * 8, hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, pic_spin_wait(ctx->stream));
    return PIC_OK;
}

// the residual of the linear problem (no Boltzmann term)
static SorArgs plain_args(const Stencil& s) {
    SorArgs a;
    a.omega = 1;
    a.ieps = 1.0 / EPS0;
    a.itwos = 1.0 / s.twos;
    a.phi0 = 0;
    a.n0 = 0;
    a.inv_te0 = 1;
    return a;
}

static int solve_learned(pic_ctx* ctx, const pic_solver_params* p, pic_solver_stats* st) {
    int rc = learned_check(ctx);
    if (rc) return rc;
    const GridC& g = ctx->g;
    size_t nn = (size_t)g.nn;
    Stencil s = make_stencil(g, false);
    const int max_fcg = std::min(p->max_it, 500);
    // after the first check, every iteration: an iteration costs far more
    // than a check (a polled wait on a few bytes)
    const int check = p->check_every > 0 ? p->check_every : 1;
    const size_t hslots = (size_t)(max_fcg + 2) * SSIZE;  // striped history, per quantity
    rc = pic_ensure_work(ctx, 8 * nn * sizeof(double) + 4 * hslots * sizeof(double));
    if (rc) return rc;
    double* r = ctx->work;
    float* z = reinterpret_cast<float*>(r + nn);  // the preconditioned residual, FP32
    double* p0 = r + 2 * nn;
    double* p1 = p0 + nn;
    double* q = p1 + nn;
    float* ftmp = reinterpret_cast<float*>(q + nn);  // four FP32 vectors: 2 nn doubles
    FcgHist h;
    h.rz = q + 3 * nn;  // past the FP32 vectors, inside the 8 nn doubles
    h.pq = h.rz + hslots;
    h.zq = h.pq + hslots;
    h.rr = h.zq + hslots;
    double* sc = ctx->scratch + LEARNED_WS + 5 * SSIZE;  // three striped scalars
    unsigned nb = stencil_blocks(s.nn), nbr = stencil_blocks_red(s.nn);
    auto precond = [&](const double* rin, float* zout, const double* qv, double* zq, double* rz) {
        return learned_precond(ctx, s, rin, zout, ftmp, qv, zq, rz);
    };
    // pinned host copies: [0, 3 S) initial sums, [3 S, 6 S) final FP64
    // residual sums, then the striped r.r history (S = SSIZE doubles)
    rc = pic_ensure_pin(ctx, (6 * SSIZE + hslots) * sizeof(double));
    if (rc) return rc;
    double* hp = ctx->pin;
    double* hrr = hp + 6 * SSIZE;
    PIC_HIP(ctx, hipMemsetAsync(sc, 0, 3 * SSIZE * sizeof(double), ctx->stream));
    k_fcg_resid<<<nb, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, ctx->phi, ctx->rho, 1.0 / EPS0, r, sc);
    PIC_HIP(ctx, hipMemcpyAsync(hp, sc, 3 * SSIZE * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    double bnorm2 = 1.0, target2 = 0.0;
    bool have_b = false;
    auto read_initial = [&]() {  // after a wait that covers the copy above
        const double rr0 = shost(hp), bb = shost(hp + SSIZE);
        bnorm2 = bb > 0 ? bb : 1.0;
        target2 = (p->tol_kind == PIC_TOL_REF_RMS) ? p->tol * p->tol * (double)nn : p->tol * p->tol * bnorm2;
        have_b = true;
        return rr0;
    };
    // A solve that follows another learned solve on this context (a time
    // loop) does not wait for the initial residual: it starts iterating and
    // reads it with the first check, since such a start is almost never
    // converged already (if it is, one iteration runs and the checks below
    // still decide convergence).  Otherwise the host waits for it here.
    const int prev = ctx->learned_last_its;
    int it = 0;
    bool conv = false;
    if (prev == 0) {
        PIC_HIP(ctx, pic_spin_wait(ctx->stream));
        conv = read_initial() <= target2;
    }
    bool have_final = false;  // FP64 residual of the final phi already copied to hp + 3 S
    if (!conv) {
        PIC_HIP(ctx, hipMemsetAsync(h.rz, 0, 4 * hslots * sizeof(double), ctx->stream));
        rc = precond(r, z, nullptr, nullptr, h.rz);  // rz[0] = z.r
        if (rc) return rc;
        double* pold = p0;
        double* pnew = p1;
        int last = 0;
        // the host reads the residual history first at 3/4 of the previous
        // solve's iteration count (a time loop's solves are alike; a later
        // first check would overshoot whenever a solve needs fewer
        // iterations than the last), then every `check` iterations
        int next_check = std::max(check, std::min(3 * prev / 4, max_fcg));
        for (int k = 0; k < max_fcg; k++) {
            {
                PicProfScope ps(ctx, PIC_K_PCG);
                if (s.ni % 4 == 0)
                    k_fcg_pq4<<<group4_blocks(s), BLOCK, 0, ctx->stream>>>(s, ctx->ntype, z, pold, pnew, q, h, k,
                                                                          nullptr);
                else
                    k_fcg_pq<<<nbr, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, z, pold, pnew, q, h, k, nullptr);
                k_fcg_update<<<nb, BLOCK, 0, ctx->stream>>>(s.nn, h, k, ctx->phi, pnew, r, q);
            }
            PIC_HIP(ctx, hipGetLastError());
            it = k + 1;
            if (it >= next_check || it == max_fcg) {
                next_check = it + check;
                PIC_HIP(ctx, hipMemcpyAsync(hrr + (size_t)last * SSIZE, h.rr + (size_t)last * SSIZE,
                                            (size_t)(it - last) * SSIZE * sizeof(double),
                                            hipMemcpyDeviceToHost, ctx->stream));
                // where convergence is expected (the previous solve's count
                // or one more), the final FP64 residual is queued with the
                // check, saving a second wait
                const bool expect = prev > 0 && it >= prev && it <= prev + 1;
                if (expect) {
                    PIC_HIP(ctx, hipMemsetAsync(sc, 0, 3 * SSIZE * sizeof(double), ctx->stream));
                    k_resid<<<nb, BLOCK, 0, ctx->stream>>>(s, plain_args(s), ctx->ntype, ctx->phi, ctx->rho, sc);
                    PIC_HIP(ctx, hipGetLastError());
                    PIC_HIP(ctx, hipMemcpyAsync(hp + 3 * SSIZE, sc, 3 * SSIZE * sizeof(double),
                                                hipMemcpyDeviceToHost, ctx->stream));
                }
                PIC_HIP(ctx, pic_spin_wait(ctx->stream));
                if (!have_b) read_initial();
                for (int j = last; j < it; j++) {
                    const double rj = shost(hrr + (size_t)j * SSIZE);
                    if (!(rj == rj)) return pic_fail(ctx, PIC_ERR_NUMERIC, "learned PCG produced NaN");
                    if (rj <= target2) conv = true;
                }
                last = it;
                if (conv) {
                    have_final = expect;
                    break;
                }
            }
            if (it == max_fcg) break;
            rc = precond(r, z, q, h.zq + (size_t)k * SSIZE, h.rz + (size_t)(k + 1) * SSIZE);
            if (rc) return rc;
            std::swap(pold, pnew);
        }
    }
    double rn[3];
    if (conv) {
        if (have_final) {
            for (int q2 = 0; q2 < 3; q2++) rn[q2] = shost(hp + (3 + q2) * SSIZE);
        } else {
            rc = residual_norms(ctx, s, plain_args(s), rn);
            if (rc) return rc;
        }
        // the recursive residual history decided convergence; the FP64
        // residual of the final phi is the judge (CG residual norms are not
        // monotone, and the recursion drifts from the true residual)
        if (!(crit_value(p, rn, (double)nn) <= p->tol)) conv = false;
    }
    if (!conv) {  // plain PCG is the fallback, from the current phi
        pic_solver_params q2 = *p;
        q2.method = PIC_SOLVER_PCG;
        pic_solver_stats s2;
        rc = solve_pcg(ctx, &q2, &s2, false);
        *st = s2;
        st->iterations += it;
        return rc;
    }
    st->converged = 1;
    st->iterations = it;
    ctx->learned_last_its = it;
    st->newton_iterations = 0;
    st->ref_l2 = std::sqrt(rn[0] / (double)nn);
    st->rel_residual = rn[2] > 0 ? std::sqrt(rn[1] / rn[2]) : std::sqrt(rn[1]);
    return PIC_OK;
}

// Newton-Boltzmann (PotentialSolver::solveNRPCG, PotentialSolver.cpp:178-241)
// with the learned preconditioner inside: per Newton step, F and the
// Jacobian shift P from the FP64 phi (k_F), then flexible CG on
// J y = F, J = A - diag(P), preconditioned by the two-level cycle whose
// smoothing sweeps, residual restriction and line search carry diag(P)
// (the network's coarse solve approximates A^-1 and the A-norm optimal step
// becomes the J-norm optimal one); phi -= y.  Converged when the FP64
// residual of the nonlinear problem over the regular rows is within the
// tolerance (PIC_TOL_REL), or, with PIC_TOL_REF_RMS, when RMS(y) < 1e-3 as
// in the reference (:229-233).  Each inner solve reduces its residual by
// 1e-2 (inexact Newton) or to below the tolerance.  Plain Newton-PCG is the
// fallback if the learned iteration stalls; the final residual is FP64.
static int solve_newton_learned(pic_ctx* ctx, const pic_solver_params* p, pic_solver_stats* st) {
    int rc = learned_check(ctx);
    if (rc) return rc;
    const GridC& g = ctx->g;
    const size_t nn = (size_t)g.nn;
    Stencil s = make_stencil(g, false);
    const int max_fcg = std::min(p->max_it, 1000);
    const size_t hslots = (size_t)(max_fcg + 2) * SSIZE;
    // r (F), z, p0, p1, q, y, P (FP64), three FP32 vectors, four histories
    rc = pic_ensure_work(ctx, 9 * nn * sizeof(double) + 4 * hslots * sizeof(double));
    if (rc) return rc;
    double* r = ctx->work;
    float* z = reinterpret_cast<float*>(r + nn);  // the preconditioned residual, FP32
    double* p0 = r + 2 * nn;
    double* p1 = p0 + nn;
    double* q = p1 + nn;
    double* y = q + nn;
    double* P = y + nn;
    float* ftmp = reinterpret_cast<float*>(P + nn);  // four FP32 vectors: 2 nn doubles
    FcgHist h;
    h.rz = P + nn + 2 * nn;
    h.pq = h.rz + hslots;
    h.zq = h.pq + hslots;
    h.rr = h.zq + hslots;
    rc = pic_ensure_pin(ctx, (4 + hslots) * sizeof(double));
    if (rc) return rc;
    double* hf = ctx->pin;       // the Newton residual sums
    double* hrr = ctx->pin + 4;  // the inner residual history (striped)
    NewtonArgs na;
    na.ieps = 1.0 / EPS0;
    na.phi0 = p->phi0;
    na.n0 = p->n0;
    na.inv_te0 = 1.0 / (p->Te0 != 0 ? p->Te0 : 1.0);
    unsigned nb = stencil_blocks(s.nn), nbr = stencil_blocks_red(s.nn);
    int total_it = 0, newton = 0;
    bool conv = false, stalled = false;
    for (newton = 0; newton < 20; newton++) {  // NR_MAX_IT, PotentialSolver.cpp:179
        PIC_HIP(ctx, hipMemsetAsync(ctx->scratch, 0, 4 * sizeof(double), ctx->stream));
        k_F<double><<<nbr, BLOCK, 0, ctx->stream>>>(s, na, ctx->ntype, ctx->phi, ctx->rho, r, P, ctx->scratch);
        PIC_HIP(ctx, hipGetLastError());
        PIC_HIP(ctx, hipMemcpyAsync(hf, ctx->scratch, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        PIC_HIP(ctx, pic_spin_wait(ctx->stream));
        const double bb = hf[2] > 0 ? hf[2] : 1.0;
        const double target2 = p->tol_kind == PIC_TOL_REF_RMS ? p->tol * p->tol * (double)nn
                                                              : p->tol * p->tol * bb;
        if (p->tol_kind == PIC_TOL_REL && hf[1] <= target2) {
            conv = true;
            break;
        }
        if (!(hf[0] == hf[0])) return pic_fail(ctx, PIC_ERR_NUMERIC, "Newton residual is NaN");
        // inexact Newton (Eisenstat-Walker style, constant forcing 1e-2):
        // each inner solve reduces its residual 100-fold, or to below the
        // tolerance; a learned iteration costs several PCG iterations, so
        // more, cheaper Newton steps pay
        const double inner2 = std::max(0.25 * target2, hf[0] * 1e-4);
        // flexible CG on J y = F from y = 0 (r = F; F vanishes off the regular rows)
        PIC_HIP(ctx, hipMemsetAsync(y, 0, nn * sizeof(double), ctx->stream));
        PIC_HIP(ctx, hipMemsetAsync(h.rz, 0, 4 * hslots * sizeof(double), ctx->stream));
        rc = learned_precond(ctx, s, r, z, ftmp, nullptr, nullptr, h.rz, P);
        if (rc) return rc;
        double* pold = p0;
        double* pnew = p1;
        int it = 0;
        bool inner_conv = false;
        for (int k = 0; k < max_fcg; k++) {
            {
                PicProfScope ps(ctx, PIC_K_PCG);
                if (s.ni % 4 == 0)
                    k_fcg_pq4<<<group4_blocks(s), BLOCK, 0, ctx->stream>>>(s, ctx->ntype, z, pold, pnew, q, h, k, P);
                else
                    k_fcg_pq<<<nbr, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, z, pold, pnew, q, h, k, P);
                k_fcg_update<<<nb, BLOCK, 0, ctx->stream>>>(s.nn, h, k, y, pnew, r, q);
            }
            PIC_HIP(ctx, hipGetLastError());
            it = k + 1;
            PIC_HIP(ctx, hipMemcpyAsync(hrr + (size_t)k * SSIZE, h.rr + (size_t)k * SSIZE, SSIZE * sizeof(double),
                                        hipMemcpyDeviceToHost, ctx->stream));
            PIC_HIP(ctx, pic_spin_wait(ctx->stream));
            const double rk = shost(hrr + (size_t)k * SSIZE);
            if (!(rk == rk)) return pic_fail(ctx, PIC_ERR_NUMERIC, "learned Newton inner solve produced NaN");
            if (rk <= inner2) {
                inner_conv = true;
                break;
            }
            if (it == max_fcg) break;
            rc = learned_precond(ctx, s, r, z, ftmp, q, h.zq + (size_t)k * SSIZE,
                                 h.rz + (size_t)(k + 1) * SSIZE, P);
            if (rc) return rc;
            std::swap(pold, pnew);
        }
        total_it += it;
        if (!inner_conv) {
            stalled = true;
            break;
        }
        PIC_HIP(ctx, hipMemsetAsync(ctx->scratch, 0, sizeof(double), ctx->stream));
        k_update_phi<double><<<nbr, BLOCK, 0, ctx->stream>>>(s, ctx->ntype, ctx->phi, y, ctx->scratch);
        PIC_HIP(ctx, hipGetLastError());
        double yy;
        PIC_HIP(ctx, hipMemcpyAsync(hf + 3, ctx->scratch, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        PIC_HIP(ctx, pic_spin_wait(ctx->stream));
        yy = hf[3];
        if (p->tol_kind == PIC_TOL_REF_RMS && std::sqrt(yy / (double)nn) < 1e-3) {  // :229-233
            conv = true;
            newton++;
            break;
        }
    }
    if (stalled || !conv) {  // plain Newton-PCG from the current phi is the fallback
        pic_solver_params q2 = *p;
        q2.method = PIC_SOLVER_PCG;
        pic_solver_stats s2;
        rc = solve_pcg(ctx, &q2, &s2, false);
        *st = s2;
        st->iterations += total_it;
        st->newton_iterations += newton;
        return rc;
    }
    SorArgs a;
    a.omega = 1;
    a.ieps = 1.0 / EPS0;
    a.itwos = 1.0 / s.twos;
    a.phi0 = p->phi0;
    a.n0 = p->n0;
    a.inv_te0 = na.inv_te0;
    double rn[3];
    rc = residual_norms(ctx, s, a, rn);
    if (rc) return rc;
    st->converged = 1;
    st->iterations = total_it;
    st->newton_iterations = newton;
    st->ref_l2 = std::sqrt(rn[0] / (double)nn);
    st->rel_residual = rn[2] > 0 ? std::sqrt(rn[1] / rn[2]) : std::sqrt(rn[1]);
    return PIC_OK;
}

// plain red-black SOR to a relative tolerance (used to make training pairs)
int pic_poisson_internal_sor(pic_ctx* ctx, double tol, int max_it, pic_solver_stats* st) {
    pic_solver_params p{};
    p.method = PIC_SOLVER_SOR_RB;
    p.tol_kind = PIC_TOL_REL;
    p.tol = tol;
    p.max_it = max_it;
    int nmax = std::max(ctx->g.ni, std::max(ctx->g.nj, ctx->g.nk));
    p.omega = 2.0 / (1.0 + std::sin(3.141592653589793 / (nmax - 1)));
    p.check_every = 25;
    int rc = pic_ensure_ntype(ctx);
    if (rc) return rc;
    return solve_sor(ctx, &p, st);
}

extern "C" int pic_poisson(pic_ctx* ctx, const pic_solver_params* p, pic_solver_stats* st) {
    if (!ctx || !p) return PIC_ERR_ARG;
    if (!(p->tol > 0) || p->max_it <= 0) return pic_fail(ctx, PIC_ERR_ARG, "bad solver params");
    int rc = pic_ensure_ntype(ctx);
    if (rc) return rc;
    pic_solver_stats local;
    if (!st) st = &local;
    *st = pic_solver_stats{};
    for (auto& e : ctx->solve_ev)
        if (!e) PIC_HIP(ctx, hipEventCreate(&e));
    hipEvent_t e0 = ctx->solve_ev[0], e1 = ctx->solve_ev[1];
    PIC_HIP(ctx, hipEventRecord(e0, ctx->stream));
    if (p->flags & PIC_SOLVE_EXTRAPOLATE) {
        const size_t nn = (size_t)ctx->g.nn;
        if (!ctx->phi_prev) PIC_HIP(ctx, hipMalloc(&ctx->phi_prev, nn * sizeof(double)));
        if (ctx->phi_prev_valid) {
            k_extrapolate<<<stencil_blocks((long long)nn), BLOCK, 0, ctx->stream>>>((long long)nn, ctx->phi,
                                                                                    ctx->phi_prev);
            PIC_HIP(ctx, hipGetLastError());
        } else {
            PIC_HIP(ctx, hipMemcpyAsync(ctx->phi_prev, ctx->phi, nn * sizeof(double), hipMemcpyDeviceToDevice,
                                        ctx->stream));
            ctx->phi_prev_valid = true;
        }
    }
    switch (p->method) {
        case PIC_SOLVER_SOR_RB: rc = solve_sor(ctx, p, st); break;
        case PIC_SOLVER_PCG: rc = solve_pcg(ctx, p, st, false); break;
        case PIC_SOLVER_PCG_MIXED: rc = solve_pcg(ctx, p, st, true); break;
        case PIC_SOLVER_LEARNED:
            rc = p->n0 != 0.0 ? solve_newton_learned(ctx, p, st) : solve_learned(ctx, p, st);
            break;
        default: rc = pic_fail(ctx, PIC_ERR_ARG, "unknown solver method");
    }
    (void)hipEventRecord(e1, ctx->stream);
    (void)pic_spin_event(e1);  // polled: the solvers' last wait has usually drained the stream
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    st->seconds = ms * 1e-3;
    return rc;
}
