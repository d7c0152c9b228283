// pic_particles.hip -- particle kernels of the gfx950 PIC library:
//   k_push     fused CIC gather + kick + drift + absorb + counting-sort
//              histogram (Species.cpp:356-399, Field.h:201-232)
//   k_hist     histogram of unchanged particles (Species.cpp:905-915)
//   k_scan_*   exclusive scan of the per-cell counts -> cell offsets
//   k_scatter  stable-per-wave scatter into the second buffer (compacts)
//   k_push_scatter  second pass of pic_push_sort: push again and write each
//              particle to its sorted position
//   k_dep_cells, k_dep_nodes  number density in two passes without atomics:
//              per-cell corner sums, then a node gather (Species.cpp:401-416,
//              Field.h:157-199)
//   k_mom_cells, k_mom_nodes  the velocity moments the same way
//              (Species.cpp:767-776)
//   k_rho      charge density (World.cpp:193-200)
//   k_load     Philox loader (Species.cpp:564-599, :420-434, :855-869)
#include <algorithm>
#include <cmath>

#include "pic_internal.h"

// Every floating-point operation in this file is rounded on its own (no
// fused multiply-add contraction): the two passes of pic_push_sort compute
// each particle's new cell from the same inputs and must agree bit for bit.
#pragma clang fp contract(off)

using namespace pic;

namespace {

constexpr int BLOCK = 256;

// Wave-aggregated counting-sort histogram.  Lanes of a wave hold
// consecutive particles; on cell-sorted input they form long runs of equal
// keys, so only the head of each run issues the atomic.  Every lane gets its
// rank inside the bucket.  All lanes of the wave must call this.
__device__ inline uint32_t run_histogram(uint32_t key, bool valid, uint32_t* __restrict__ count) {
    const int lane = threadIdx.x & 63;
    uint32_t prev = __shfl_up(key, 1, 64);
    bool head = (lane == 0) || (key != prev);
    unsigned long long heads = __ballot(head);
    unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    int my_head = 63 - __clzll(heads & upto);
    unsigned long long above = heads & ~upto;
    int next_head = above ? (__ffsll((unsigned long long)above) - 1) : 64;
    uint32_t base = 0;
    if (head && valid) base = atomicAdd(&count[key], (uint32_t)(next_head - lane));
    base = __shfl(base, my_head, 64);
    return base + (uint32_t)(lane - my_head);
}

// The runs of equal keys in a wave's 64 consecutive particles: the lane's
// run index (heads counted up to and including its run's head) and the
// lane of that head.  All lanes of the wave must call this.
__device__ inline void key_runs(uint32_t key, int& run, int& head_lane, int& run_len, bool& head) {
    const int lane = threadIdx.x & 63;
    uint32_t prev = __shfl_up(key, 1, 64);
    head = (lane == 0) || (key != prev);
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    head_lane = 63 - __clzll(heads & upto);
    run = __popcll(heads & upto) - 1;
    const unsigned long long above = heads & ~upto;
    run_len = (above ? (__ffsll((unsigned long long)above) - 1) : 64) - lane;  // valid in head lanes
}

// The histogram pass of pic_push_sort: as run_histogram, and each run's
// first slot in its bucket is kept at bases[run] (bases: the 64 entries of
// this wave's group of 64 particles), so that the second pass, which sees
// the same groups and keys, finds every particle's slot without atomics.
__device__ inline void run_bases(uint32_t key, bool valid, uint32_t* __restrict__ count,
                                 uint32_t* __restrict__ bases) {
    int run, head_lane, len;
    bool head;
    key_runs(key, run, head_lane, len, head);
    if (head && valid) bases[run] = atomicAdd(&count[key], (uint32_t)len);
}

// What a particle that ends its drift inside an object does (PIC_SURF_*);
// see surface_path.
struct SurfArgs {
    int mode;
    int emit;          // this pass writes emitted particles (the pass that writes the push)
    double dt;
    double vth_wall;   // sqrt(2 k 300 K / m) of the pushed species (Species.cpp:836, :856)
    double mpw0_t;     // weight of an emitted particle (the target's mpw0, Species.cpp:225)
    uint32_t seed, step;
    SpeciesBuf tgt;    // emission target (PIC_SURF_RECYCLE)
    uint32_t* tgt_n;   // its counts: [0] live count, [2] emissions that did not fit
    uint32_t tgt_cap;
};

struct PushArgs {
    float qmdt;    // dt * q / m   (Species.cpp:372)
    float dtx[3];  // dt / dx      (drift in cells)
    int kick;
    // the field's per-component |E| maxima, for the histogram pass's test of
    // particles that cannot leave their cell; null: every particle is pushed
    const float* emax;
    SurfArgs s;
};

// A particle whose drift cannot take it out of its cell, whatever E it
// gathers: for every axis |f' - f| <= (|v| + |qm dt| max|E|) dt/dx (with
// margins for the FP32 rounding of the kick, the interpolation and the
// drift) leaves f' strictly inside (0, 1).  Its new cell is its old one, so
// the histogram pass needs no gather for it; the pass that writes pushes
// every particle and finds the same cell.
__device__ inline bool stays_in_cell(const float (&f)[3], const float (&v)[3], const float (&kick)[3],
                                     const float (&dtx)[3]) {
    bool in = true;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const float m = (fabsf(v[d]) + kick[d]) * fabsf(dtx[d]) * 1.00001f + 1e-6f;
        in = in && (f[d] > m) && (f[d] + m < 1.0f);  // false for NaN
    }
    return in;
}

// 32-bit node offsets (grids hold fewer than 2^31 nodes)
__device__ inline float3 gather_e(const GridC& g, const float4* __restrict__ E, int i, int j, int k,
                                  float fx, float fy, float fz) {
    const unsigned sj = (unsigned)g.ni, sk = (unsigned)g.ni * (unsigned)g.nj;
    const unsigned u = (unsigned)k * sk + (unsigned)j * sj + (unsigned)i;
    // four node pairs (i, i+1): one address each, the +1 is an immediate offset
    const float4* p00 = E + u;
    const float4* p01 = p00 + sk;
    const float4* p10 = p00 + sj;
    const float4* p11 = p10 + sk;
    float4 e000 = p00[0], e001 = p01[0], e010 = p10[0], e011 = p11[0];
    float4 e100 = p00[1], e101 = p01[1], e110 = p10[1], e111 = p11[1];
    float gx = 1.f - fx, gy = 1.f - fy, gz = 1.f - fz;
    float a = gx * gy, b = gx * fy, c = fx * gy, d = fx * fy;
    float w000 = a * gz, w001 = a * fz, w010 = b * gz, w011 = b * fz;
    float w100 = c * gz, w101 = c * fz, w110 = d * gz, w111 = d * fz;
    float3 r;
    r.x = e000.x * w000 + e001.x * w001 + e010.x * w010 + e011.x * w011 + e100.x * w100 +
          e101.x * w101 + e110.x * w110 + e111.x * w111;
    r.y = e000.y * w000 + e001.y * w001 + e010.y * w010 + e011.y * w011 + e100.y * w100 +
          e101.y * w101 + e110.y * w110 + e111.y * w111;
    r.z = e000.z * w000 + e001.z * w001 + e010.z * w010 + e011.z * w011 + e100.z * w100 +
          e101.z * w101 + e110.z * w110 + e111.z * w111;
    return r;
}

// move the in-cell offset by dcell cells; returns the new cell coordinate
__device__ inline int relocate(float& f, int i) {
    float o = floorf(f);
    f -= o;
    if (f >= 1.0f) {  // f was a tiny negative number
        f = 0.0f;
        o += 1.0f;
    }
    o = fminf(fmaxf(o, -4096.0f), 4096.0f);
    return i + (int)o;
}

// Rectangle::inObject (Object.cpp:231-238): inclusive box, decided in FP64
// on the exact logical coordinate i + f
__device__ inline bool in_rect(const GridC& g, int i, int j, int k, float fx, float fy, float fz) {
    const double l[3] = {(double)i + (double)fx, (double)j + (double)fy, (double)k + (double)fz};
    for (int r = 0; r < g.nrect; r++) {
        bool in = true;
#pragma unroll
        for (int d = 0; d < 3; d++)
            if (l[d] < g.rlo[r][d] || l[d] > g.rhi[r][d]) in = false;
        if (in) return true;
    }
    return false;
}

// ---- object surfaces (Species.cpp:170-256), FP64, entered only by
// particles of a reflecting or recycling species that end inside an object ----

// first rectangle holding logical point l (Rectangle::inObject, inclusive;
// World::inObject takes the first, World.cpp:293-301), -1 if none
__device__ inline int rect_index(const GridC& g, const double* l) {
    for (int r = 0; r < g.nrect; r++) {
        bool in = true;
#pragma unroll
        for (int d = 0; d < 3; d++)
            if (l[d] < g.rlo[r][d] || l[d] > g.rhi[r][d]) in = false;
        if (in) return r;
    }
    return -1;
}

// Rectangle::lineIntersect (Object.cpp:271-308) in logical coordinates (an
// affine map per axis leaves the entry time unchanged): slab entry time,
// the entry side's outward normal, summed and normalised with every side
// whose time is within 1e-6; pos = the entry point
__device__ inline double rect_entry(const GridC& g, int r, const double* x1, const double* x2,
                                    double* pos, double* n) {
    double tmin[3], tmax[3];
    int side[3] = {0, 1, 2};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double A = x2[d] - x1[d];
        tmin[d] = (g.rlo[r][d] - x1[d]) / A;
        tmax[d] = (g.rhi[r][d] - x1[d]) / A;
        if (tmin[d] > tmax[d]) {
            const double t = tmin[d];
            tmin[d] = tmax[d];
            tmax[d] = t;
            side[d] += 3;
        }
    }
    double te = tmin[0];
    int j = 0, k = 0, t = 0;
    for (int i = 0; i < 2; i++) {
        if (te < tmin[i + 1]) {
            te = tmin[i + 1];
            j = i + 1;
            k = 0;
        } else if (fabs(te - tmin[i + 1]) < 1e-6) {
            k++;
            t = i + 1;
        }
    }
    n[0] = n[1] = n[2] = 0.0;
    auto add = [&](int sd) { n[sd % 3] += sd < 3 ? -1.0 : 1.0; };  // Object.cpp:310-316
    add(side[j]);
    if (k != 0) {
        if (k < 2) {
            add(side[t]);
        } else {
            add(side[2]);
            add(side[1]);
        }
        const double inv = 1.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
#pragma unroll
        for (int d = 0; d < 3; d++) n[d] *= inv;
    }
#pragma unroll
    for (int d = 0; d < 3; d++) pos[d] = x1[d] + te * (x2[d] - x1[d]);
    return te;
}

__device__ inline void cross3d(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// sampleVth(300 K) then sampleReflectedVelocity (Species.cpp:855-861,
// :835-853): full accommodation, the direction built as in the source from
// t1 = n x e_x (n x e_y when n . e_x == 0) and t2 = n x t1, not normalised,
// sin(psi) t2 without sin(theta).  Draws s0 .. s0 + 10.
__device__ inline void reflect_velocity(double vth0, double vmag1, const double* n, Draws& dr, int s0,
                                        double* out) {
    double v[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double u1 = dr(s0 + 3 * c), u2 = dr(s0 + 3 * c + 1), u3 = dr(s0 + 3 * c + 2);
        v[c] = vth0 * (u1 + u2 + u3 - 1.5);
    }
    const double v_th = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double vmag2 = vmag1 + 1.0 * (v_th - vmag1);
    const double sin_theta = dr(s0 + 9);
    const double cos_theta = sqrt(1 - sin_theta * sin_theta);
    const double psi = 2 * PI_REF * dr(s0 + 10);
    const double ex[3] = {1, 0, 0}, ey[3] = {0, 1, 0};
    double t1[3], t2[3];
    if (n[0] * 1.0 + n[1] * 0.0 + n[2] * 0.0 != 0)
        cross3d(n, ex, t1);
    else
        cross3d(n, ey, t1);
    cross3d(n, t1, t2);
    const double a = sin_theta * cos(psi), b = sin(psi);
#pragma unroll
    for (int d = 0; d < 3; d++) out[d] = ((t1[d] * a + t2[d] * b) + n[d] * cos_theta) * vmag2;
}

// The drift of a particle whose FP32 drift ended inside an object, redone
// in FP64 from its stored position with the kicked velocity v, as the
// reference's bounce loop (Species.cpp:194-249): leaving the box removes;
// ending in an object moves the particle to the entry point, where a
// reflecting species takes a diffusely reflected velocity and drifts on for
// the remaining fraction of the step (at most 20 bounces), and a recycling
// species is removed after emitting int(w / mpw0_t + U) particles of the
// target there (when s.emit).  Emitted particles at the entry point are
// kept (the source's addParticle rejects those its rounding leaves inside
// the object, Species.cpp:424).  Both passes of pic_push_sort run this with
// the same inputs and draws and so agree.  Updates f and v; returns the
// new cell word or DEAD.
__device__ inline uint32_t surface_path(const GridC& g, const SurfArgs& s, uint32_t c,
                                              const SpeciesBuf& b, size_t p, float (&f)[3],
                                              float (&v)[3]) {
    int ijk[3];
    unpack_cell(c, ijk[0], ijk[1], ijk[2]);
    double l[3], vel[3], step[3];
    const int lim[3] = {g.ci, g.cj, g.ck};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        l[d] = (double)ijk[d] + (double)b.f[d][p];  // the position before this push
        vel[d] = (double)v[d];
        step[d] = s.dt * g.inv_dx[d];
    }
    double t_rem = 1.0;
    int nb = 0;
    while (t_rem > 0) {
        if (++nb > 20) return DEAD;
        double old[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            old[d] = l[d];
            l[d] += vel[d] * t_rem * step[d];
        }
        bool inb = true;
#pragma unroll
        for (int d = 0; d < 3; d++)
            if (!(l[d] >= 0.0 && l[d] < (double)lim[d])) inb = false;
        if (!inb) return DEAD;
        const int r = rect_index(g, l);
        if (r < 0) break;
        double nrm[3];
        const double tp = rect_entry(g, r, old, l, l, nrm);
        const double vm = sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
        Draws dr;
        dr.init((uint32_t)p, (uint32_t)nb, s.seed, s.step, TAG_SURF);
        if (s.mode == PIC_SURF_REFLECT) {
            reflect_velocity(s.vth_wall, vm, nrm, dr, 0, vel);
            t_rem *= (1 - tp);
            continue;
        }
        if (s.emit) {  // PIC_SURF_RECYCLE
            const int mp = (int)((double)b.w[p] / s.mpw0_t + dr(0));
            int cc[3];
            float fe[3];
            bool inside = true;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                if (!(l[d] >= 0.0 && l[d] < (double)lim[d])) inside = false;
                cc[d] = (int)l[d];
                const float fr = (float)(l[d] - (double)cc[d]);
                fe[d] = fr < 1.0f ? fr : 0.99999994f;
            }
            for (int m = 0; inside && m < mp; m++) {
                double vn[3];
                reflect_velocity(s.vth_wall, vm, nrm, dr, 1 + 11 * m, vn);
                const uint32_t idx = atomicAdd(s.tgt_n, 1u);
                if (idx >= s.tgt_cap) {
                    atomicAdd(s.tgt_n + 2, 1u);
                    continue;
                }
                s.tgt.cell[idx] = pack_cell(cc[0], cc[1], cc[2]);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    s.tgt.f[d][idx] = fe[d];
                    s.tgt.v[d][idx] = (float)vn[d];
                }
                s.tgt.w[idx] = (float)s.mpw0_t;
            }
        }
        return DEAD;
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        ijk[d] = (int)l[d];
        const float fr = (float)(l[d] - (double)ijk[d]);
        f[d] = fr < 1.0f ? fr : 0.99999994f;
        v[d] = (float)vel[d];
    }
    return pack_cell(ijk[0], ijk[1], ijk[2]);
}

// One particle's push (Species.cpp:356-399): CIC gather of E at the old
// position, kick, drift, relocation into the new cell, removal at the walls
// (World::inBounds, World.cpp:201-205); a particle ending inside an object
// is removed or, for a reflecting or recycling species, handed to
// surface_path (b and p locate its stored position).  Updates f and v;
// returns the new cell word, DEAD when the particle is removed.
template <bool SURF>
__device__ inline uint32_t advance(const GridC& g, const float4* __restrict__ E, const PushArgs& a,
                                   uint32_t c, float (&f)[3], float (&v)[3], const SpeciesBuf& b,
                                   size_t p) {
    int i, j, k;
    unpack_cell(c, i, j, k);
    if (a.kick) {
        float3 e = gather_e(g, E, i, j, k, f[0], f[1], f[2]);
        v[0] += e.x * a.qmdt;
        v[1] += e.y * a.qmdt;
        v[2] += e.z * a.qmdt;
    }
    f[0] += v[0] * a.dtx[0];
    f[1] += v[1] * a.dtx[1];
    f[2] += v[2] * a.dtx[2];
    bool ok = (f[0] == f[0]) && (f[1] == f[1]) && (f[2] == f[2]);
    if (ok) {
        i = relocate(f[0], i);
        j = relocate(f[1], j);
        k = relocate(f[2], k);
        ok = i >= 0 && i < g.ci && j >= 0 && j < g.cj && k >= 0 && k < g.ck;
        if (ok && g.nrect && in_rect(g, i, j, k, f[0], f[1], f[2])) {
            if (!SURF) return DEAD;  // the kernels without the surface path serve absorbing species
            return surface_path(g, a.s, c, b, p, f, v);
        }
    }
    return ok ? pack_cell(i, j, k) : DEAD;
}

// the histogram key of a cell word: the cell index, ncells for removed
__device__ inline uint32_t cell_key(const GridC& g, uint32_t c) {
    if (c == DEAD) return (uint32_t)g.ncells;
    int i, j, k;
    unpack_cell(c, i, j, k);
    return (uint32_t)cell_index(g, i, j, k);
}

// WRITE: the push proper (positions, velocities, changed cell words and the
// ranks inside the cells are written); otherwise only the counting sort's
// histogram of the pushed particles (the first pass of pic_push_sort).
// PUSH_ILP particles per thread, BLOCK apart: every thread issues the loads
// of all its particles (cell word, offsets, velocities: no load waits on the
// cell word) before it waits for any, so twice the bytes are in flight.
constexpr int PUSH_ILP = 2;
// the histogram pass of pic_push_sort (4 per thread measured slower)
constexpr int PUSH_SORT_ILP = 2;
template <bool WRITE, int ILP, bool SURF>
__global__ void __launch_bounds__(BLOCK)
    k_push(GridC g, SpeciesBuf b, const float4* __restrict__ E, PushArgs a,
           uint32_t* __restrict__ d_n, uint32_t* __restrict__ count, uint32_t* __restrict__ rank) {
    const size_t p0 = xcd_wg() * (BLOCK * ILP) + threadIdx.x;
    uint32_t n = d_n[0];
    uint32_t covered = gridDim.x * BLOCK * ILP;
    if (n > covered) n = covered;
    if (p0 == 0) d_n[1] = n;
    uint32_t c[ILP];
    float f[ILP][3], v[ILP][3];
    // the largest kick per axis (histogram pass, no objects)
    const bool bound = !WRITE && !SURF && a.emax != nullptr;
    float kick[3] = {0.f, 0.f, 0.f};
    if (bound && a.kick)
#pragma unroll
        for (int d = 0; d < 3; d++) kick[d] = fabsf(a.qmdt) * a.emax[d];
#pragma unroll
    for (int q = 0; q < ILP; q++) {
        const size_t p = p0 + (size_t)q * BLOCK;
        c[q] = DEAD;
        if (p < n) {
            c[q] = b.cell[p];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                f[q][d] = b.f[d][p];
                v[q][d] = b.v[d][p];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < ILP; q++) {
        const size_t p = p0 + (size_t)q * BLOCK;
        const bool valid = p < n;
        uint32_t key = 0xFFFFFFFFu;
        if (valid) {
            key = (uint32_t)g.ncells;
            if (c[q] != DEAD) {
                const uint32_t nc = (bound && stays_in_cell(f[q], v[q], kick, a.dtx))
                                        ? c[q]
                                        : advance<SURF>(g, E, a, c[q], f[q], v[q], b, p);
                key = cell_key(g, nc);
                if (WRITE) {
                    if (nc != c[q]) b.cell[p] = nc;  // most particles stay in their cell
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        b.f[d][p] = f[q][d];
                        b.v[d][p] = v[q][d];
                    }
                }
            }
        }
        if (WRITE) {
            const uint32_t r = run_histogram(key, valid, count);
            if (valid) rank[p] = r;
        } else {
            run_bases(key, valid, count, rank + (p & ~(size_t)63));  // the group's run bases
        }
    }
}

__global__ void __launch_bounds__(BLOCK)
    k_hist(GridC g, SpeciesBuf b, uint32_t* __restrict__ d_n, uint32_t* __restrict__ count,
           uint32_t* __restrict__ rank, uint32_t start) {  // particles [start, live count)
    size_t p = start + xcd_wg() * BLOCK + threadIdx.x;
    uint32_t n = d_n[0];
    uint32_t covered = start + gridDim.x * BLOCK;
    if (n > covered) n = covered;
    if (p == start) d_n[1] = n;
    bool valid = p < n;
    uint32_t key = 0xFFFFFFFFu;
    if (valid) {
        uint32_t c = b.cell[p];
        key = (uint32_t)g.ncells;
        if (c != DEAD) {
            int i, j, k;
            unpack_cell(c, i, j, k);
            key = (uint32_t)cell_index(g, i, j, k);
        }
    }
    uint32_t r = run_histogram(key, valid, count);
    if (valid) rank[p] = r;
}

// ---- exclusive scan of count[0..M) -> off[0..M], M = ncells + 1 ----
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[BLOCK / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; w++) {
        if (w < wv) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    if (total) *total = all;
    return before + incl - v;
}

__global__ void __launch_bounds__(BLOCK) k_scan_reduce(const uint32_t* __restrict__ in, int M,
                                                       uint32_t* __restrict__ bsum) {
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++)
        if (base + q < (size_t)M) s += in[base + q];
    uint32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// single workgroup: exclusive scan of the tile sums in place
__global__ void __launch_bounds__(BLOCK) k_scan_top(uint32_t* __restrict__ bsum, int nb) {
    int per = (nb + BLOCK - 1) / BLOCK;
    int beg = threadIdx.x * per;
    uint32_t s = 0;
    for (int q = 0; q < per; q++)
        if (beg + q < nb) s += bsum[beg + q];
    uint32_t ex = block_exclusive_scan(s, nullptr);
    for (int q = 0; q < per; q++)
        if (beg + q < nb) {
            uint32_t t = bsum[beg + q];
            bsum[beg + q] = ex;
            ex += t;
        }
}

__global__ void __launch_bounds__(BLOCK)
    k_scan_down(const uint32_t* __restrict__ in, int M, const uint32_t* __restrict__ bsum,
                uint32_t* __restrict__ out, int ncells, uint32_t* __restrict__ d_n) {
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++) {
        v[q] = (base + q < (size_t)M) ? in[base + q] : 0u;
        s += v[q];
    }
    uint32_t ex = block_exclusive_scan(s, nullptr) + bsum[blockIdx.x];
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; q++) {
        size_t idx = base + q;
        if (idx < (size_t)M) {
            out[idx] = ex;
            if (idx == (size_t)ncells && d_n) {  // particles alive after compaction
                d_n[0] = ex;
                d_n[3] = ex;  // all of them sorted: the sorted part of a two-part species
            }
            if (idx == (size_t)M - 1) out[M] = ex + v[q];
        }
        ex += v[q];
    }
}

// A workgroup moves SC_V * BLOCK consecutive particles: it reads them with
// 16-byte loads (SC_V = 4 consecutive particles per lane, every array), stages
// them in LDS, and writes them out one particle per lane so consecutive lanes
// store to consecutive destinations (the order inside a cell is the rank from
// the push's histogram).  Vector loads only where all four particles are
// below the live count; the tail is read element by element.
constexpr int SC_V = 4;
__global__ void __launch_bounds__(BLOCK)
    k_scatter(GridC g, SpeciesBuf src, SpeciesBuf dst, const uint32_t* __restrict__ rank,
              const uint32_t* __restrict__ off, const uint32_t* __restrict__ d_n, uint32_t cap,
              uint32_t start) {  // particles [start, d_n[1]) to start + off[cell] + rank
    constexpr int NP = BLOCK * SC_V;
    __shared__ uint4 st[9][BLOCK];  // cell, fx, fy, fz, vx, vy, vz, w, rank; 4 particles each
    const uint32_t n = d_n[1];
    const size_t base = (start & ~3u) + xcd_wg() * NP;  // 16-byte aligned groups of 4
    const int tid = threadIdx.x;
    const size_t p4 = base + (size_t)SC_V * tid;
    const uint32_t* arr[9] = {src.cell,
                              reinterpret_cast<const uint32_t*>(src.f[0]),
                              reinterpret_cast<const uint32_t*>(src.f[1]),
                              reinterpret_cast<const uint32_t*>(src.f[2]),
                              reinterpret_cast<const uint32_t*>(src.v[0]),
                              reinterpret_cast<const uint32_t*>(src.v[1]),
                              reinterpret_cast<const uint32_t*>(src.v[2]),
                              reinterpret_cast<const uint32_t*>(src.w),
                              rank};
    uint4 r[9];
    if (p4 >= start && p4 + SC_V <= n) {
#pragma unroll
        for (int a = 0; a < 9; a++) r[a] = *reinterpret_cast<const uint4*>(arr[a] + p4);
    } else {
#pragma unroll
        for (int a = 0; a < 9; a++) {
            uint32_t e[4] = {DEAD, DEAD, DEAD, DEAD};  // past the live count: skipped below
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (p4 + k >= start && p4 + k < n) e[k] = arr[a][p4 + k];
            r[a] = make_uint4(e[0], e[1], e[2], e[3]);
        }
    }
#pragma unroll
    for (int a = 0; a < 9; a++) st[a][tid] = r[a];
    __syncthreads();
    const uint32_t* s32[9];
#pragma unroll
    for (int a = 0; a < 9; a++) s32[a] = reinterpret_cast<const uint32_t*>(st[a]);
#pragma unroll
    for (int e = 0; e < SC_V; e++) {
        const int q = tid + BLOCK * e;
        if (base + q >= n || base + q < start) continue;
        const uint32_t c = s32[0][q];
        if (c == DEAD) continue;
        int i, j, k;
        unpack_cell(c, i, j, k);
        const uint32_t dst_i = start + off[cell_index(g, i, j, k)] + s32[8][q];
        if (dst_i >= cap) continue;  // cannot happen with a current histogram
        dst.cell[dst_i] = c;
        dst.f[0][dst_i] = __uint_as_float(s32[1][q]);
        dst.f[1][dst_i] = __uint_as_float(s32[2][q]);
        dst.f[2][dst_i] = __uint_as_float(s32[3][q]);
        dst.v[0][dst_i] = __uint_as_float(s32[4][q]);
        dst.v[1][dst_i] = __uint_as_float(s32[5][q]);
        dst.v[2][dst_i] = __uint_as_float(s32[6][q]);
        dst.w[dst_i] = __uint_as_float(s32[7][q]);
    }
}

// ---- sort of scrambled input (after a load or an upload) ----
// A counting sort writes each particle to off[cell] + rank: on input in
// random order every histogram atomic returns from a different line and
// every store lands on its own cache line, eight arrays per particle.
// Scrambled input is therefore sorted as an MSD radix sort on the cell key:
//   * level 1 over the whole array, up to P1_B buckets: per tile of PART_TILE particles a bucket
//     histogram (k_part_count), one scan over the bucket-major (bucket, tile)
//     counts, and k_part_move, which orders each tile by bucket in LDS and
//     writes every bucket's share as one contiguous run;
//   * further levels inside each bucket of the previous one, splitting it
//     into up to PART_B buckets (tiles never straddle a bucket: k_seg_tiles,
//     k_seg_count, k_seg_scan, k_seg_move), until a bucket spans at most
//     FINE_CELLS keys;
//   * k_fine_sort: one workgroup per final bucket ranks its particles by
//     cell with LDS atomics, writes the cells' offsets and moves the
//     particles through LDS with coalesced loads and stores (a bucket larger
//     than LDS takes global atomics and a scatter confined to its range).
// Every pass streams: a tile's run per bucket is PART_TILE / P1_B or more
// particles long.  The order inside a cell is that of the LDS atomics.
constexpr int PART_TILE = 4096, P1_B = 128, PART_B = 64, PART_BLOCK = 256;
constexpr int FINE_SHIFT = 9, FINE_CELLS = 1 << FINE_SHIFT;  // keys of a final bucket
constexpr int FINE_CAP = 12288;  // particles a final bucket sorts in LDS

__device__ inline uint32_t part_bucket(const GridC& g, uint32_t c, int shift) { return cell_key(g, c) >> shift; }

// One tile [base, base + cnt) ordered by bucket ((key >> shift) - bsub, in
// [0, nbk)) and written out: bucket b's run starts at start[b * stride].
__device__ inline void part_tile(const GridC& g, const SpeciesBuf& src, const SpeciesBuf& dst, size_t base,
                                 int cnt, int shift, uint32_t bsub, int nbk, const uint32_t* __restrict__ start,
                                 size_t stride) {
    __shared__ uint32_t lstart[P1_B];  // bucket counts, then the tile's local bucket starts
    __shared__ uint32_t lbase[P1_B];   // global position of the bucket's run minus its local start
    __shared__ uint16_t lpos[PART_TILE];   // tile particle q -> its position in the tile's bucket order
    __shared__ uint32_t ldest[PART_TILE];  // bucket-order position r -> global position
    __shared__ uint32_t stage[PART_TILE];  // one array at a time, in bucket order
    uint16_t* lkey = reinterpret_cast<uint16_t*>(stage);  // (set-up only, aliasing stage)
    uint16_t* lslot = lkey + PART_TILE;
    for (int b = threadIdx.x; b < nbk; b += PART_BLOCK) lstart[b] = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < cnt; q += PART_BLOCK) {
        const uint32_t k = part_bucket(g, src.cell[base + q], shift) - bsub;
        lkey[q] = (uint16_t)k;
        lslot[q] = (uint16_t)atomicAdd(&lstart[k], 1u);
    }
    __syncthreads();
    {
        const uint32_t v = threadIdx.x < nbk ? lstart[threadIdx.x] : 0u;
        const uint32_t ex = block_exclusive_scan(v, nullptr);
        if (threadIdx.x < nbk) {
            lstart[threadIdx.x] = ex;
            lbase[threadIdx.x] = start[(size_t)threadIdx.x * stride] - ex;
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < cnt; q += PART_BLOCK) {
        const uint32_t k = lkey[q], r = lstart[k] + lslot[q];
        lpos[q] = (uint16_t)r;
        ldest[r] = lbase[k] + r;
    }
    __syncthreads();
    const uint32_t* in[8] = {src.cell,
                             reinterpret_cast<const uint32_t*>(src.f[0]),
                             reinterpret_cast<const uint32_t*>(src.f[1]),
                             reinterpret_cast<const uint32_t*>(src.f[2]),
                             reinterpret_cast<const uint32_t*>(src.v[0]),
                             reinterpret_cast<const uint32_t*>(src.v[1]),
                             reinterpret_cast<const uint32_t*>(src.v[2]),
                             reinterpret_cast<const uint32_t*>(src.w)};
    uint32_t* out[8] = {dst.cell,
                        reinterpret_cast<uint32_t*>(dst.f[0]),
                        reinterpret_cast<uint32_t*>(dst.f[1]),
                        reinterpret_cast<uint32_t*>(dst.f[2]),
                        reinterpret_cast<uint32_t*>(dst.v[0]),
                        reinterpret_cast<uint32_t*>(dst.v[1]),
                        reinterpret_cast<uint32_t*>(dst.v[2]),
                        reinterpret_cast<uint32_t*>(dst.w)};
    // each thread's PER slots (q = r = threadIdx.x + e * PART_BLOCK) in
    // registers; every array's loads issued before their LDS stores
    constexpr int PER = PART_TILE / PART_BLOCK;
    uint32_t mypos[PER], mydest[PER];
#pragma unroll
    for (int e = 0; e < PER; e++) {
        const int q = threadIdx.x + e * PART_BLOCK;
        mypos[e] = q < cnt ? lpos[q] : 0u;
        mydest[e] = q < cnt ? ldest[q] : 0u;
    }
    // (the next array's loads are issued before this array's stores)
    uint32_t v[PER];
#pragma unroll
    for (int e = 0; e < PER; e++) {
        const int q = threadIdx.x + e * PART_BLOCK;
        v[e] = q < cnt ? in[0][base + q] : 0u;
    }
#pragma unroll 1
    for (int a = 0; a < 8; a++) {
#pragma unroll
        for (int e = 0; e < PER; e++)
            if (threadIdx.x + e * PART_BLOCK < cnt) stage[mypos[e]] = v[e];
        __syncthreads();
        if (a < 7) {
#pragma unroll
            for (int e = 0; e < PER; e++) {
                const int q = threadIdx.x + e * PART_BLOCK;
                v[e] = q < cnt ? in[a + 1][base + q] : 0u;
            }
        }
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int r = threadIdx.x + e * PART_BLOCK;
            if (r < cnt) out[a][mydest[e]] = stage[r];
        }
        __syncthreads();
    }
}

// level 1: tiles of the whole array [0, d_n[0])
__global__ void __launch_bounds__(PART_BLOCK)
    k_part_count(GridC g, const uint32_t* __restrict__ cell, const uint32_t* __restrict__ d_n, int shift,
                 int nbk, uint32_t ntile, uint32_t* __restrict__ pcount) {
    __shared__ uint32_t h[P1_B];
    for (int b = threadIdx.x; b < nbk; b += PART_BLOCK) h[b] = 0;
    __syncthreads();
    const size_t n = d_n[0], base = (size_t)blockIdx.x * PART_TILE;
    for (int q = threadIdx.x; q < PART_TILE; q += PART_BLOCK)
        if (base + q < n) atomicAdd(&h[part_bucket(g, cell[base + q], shift)], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < nbk; b += PART_BLOCK) pcount[(size_t)b * ntile + blockIdx.x] = h[b];
}
__global__ void __launch_bounds__(PART_BLOCK)
    k_part_move(GridC g, SpeciesBuf src, SpeciesBuf dst, const uint32_t* __restrict__ d_n, int shift, int nbk,
                uint32_t ntile, const uint32_t* __restrict__ poff) {
    const size_t n = d_n[0], base = (size_t)blockIdx.x * PART_TILE;
    const int cnt = base < n ? (int)min((size_t)PART_TILE, n - base) : 0;
    part_tile(g, src, dst, base, cnt, shift, 0u, nbk, poff + blockIdx.x, ntile);
}
// level-1 bucket bounds: bound[b] = poff[b * ntile] (b <= nbk)
__global__ void k_part_bounds(const uint32_t* __restrict__ poff, int nbk, uint32_t ntile, uint32_t* __restrict__ bound) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b <= nbk) bound[b] = poff[(size_t)b * ntile];
}

// further levels: segment s = [bound[s], bound[s + 1]) holds keys with
// key >> (shift + step) == s, split into 1 << step buckets (step <= 6);
// its tiles are numbered from tbase[s]
__global__ void __launch_bounds__(PART_BLOCK)
    k_seg_tiles(const uint32_t* __restrict__ bound, int nseg, uint32_t* __restrict__ tbase) {
    constexpr int PER = 64;  // segments per thread (nseg <= PER * PART_BLOCK)
    uint32_t v[PER], sum = 0;
    for (int e = 0; e < PER; e++) {
        const int s = threadIdx.x * PER + e;
        v[e] = s < nseg ? (bound[s + 1] - bound[s] + PART_TILE - 1) / PART_TILE : 0u;
        sum += v[e];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan(sum, &total);
    for (int e = 0; e < PER; e++) {
        const int s = threadIdx.x * PER + e;
        if (s < nseg) tbase[s] = ex;
        ex += v[e];
    }
    if (threadIdx.x == 0) tbase[nseg] = total;
}
__device__ inline bool seg_tile(const uint32_t* __restrict__ bound, const uint32_t* __restrict__ tbase, int nseg,
                                uint32_t u, int& s, size_t& base, int& cnt) {
    if (u >= tbase[nseg]) return false;
    int lo = 0, hi = nseg - 1;  // the last segment with tbase[s] <= u
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tbase[mid] <= u)
            lo = mid;
        else
            hi = mid - 1;
    }
    s = lo;
    base = (size_t)bound[s] + (size_t)(u - tbase[s]) * PART_TILE;
    cnt = (int)min((size_t)PART_TILE, (size_t)bound[s + 1] - base);
    return true;
}
__global__ void __launch_bounds__(PART_BLOCK)
    k_seg_count(GridC g, const uint32_t* __restrict__ cell, const uint32_t* __restrict__ bound,
                const uint32_t* __restrict__ tbase, int nseg, int shift, int step, uint32_t* __restrict__ cnt2) {
    __shared__ uint32_t h[PART_B];
    int s, cnt;
    size_t base;
    if (!seg_tile(bound, tbase, nseg, blockIdx.x, s, base, cnt)) return;  // (whole workgroup)
    for (int b = threadIdx.x; b < PART_B; b += PART_BLOCK) h[b] = 0;
    __syncthreads();
    const uint32_t bsub = (uint32_t)s << step;
    for (int q = threadIdx.x; q < cnt; q += PART_BLOCK) atomicAdd(&h[part_bucket(g, cell[base + q], shift) - bsub], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < PART_B; b += PART_BLOCK) cnt2[(size_t)blockIdx.x * PART_B + b] = h[b];
}
// one workgroup of PART_B threads per segment: thread b scans bucket b's
// counts over the segment's tiles (into run starts) and the buckets' totals
// give the next level's bounds
__global__ void __launch_bounds__(PART_B)
    k_seg_scan(const uint32_t* __restrict__ bound, const uint32_t* __restrict__ tbase, int nseg, int step,
               uint32_t* __restrict__ cnt2, uint32_t* __restrict__ nbound) {
    const int s = blockIdx.x, b = threadIdx.x, nsub = 1 << step;
    const uint32_t u0 = tbase[s], u1 = b < nsub ? tbase[s + 1] : tbase[s];  // lanes past nsub: none
    uint32_t run = 0;
    for (uint32_t u = u0; u < u1; u++) {
        const uint32_t c = cnt2[(size_t)u * PART_B + b];
        cnt2[(size_t)u * PART_B + b] = run;
        run += c;
    }
    uint32_t incl = run;  // wave-wide inclusive scan of the totals (PART_B = 64 lanes)
    for (int o = 1; o < PART_B; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, PART_B);
        if (b >= o) incl += t;
    }
    const uint32_t bst = bound[s] + incl - run;
    for (uint32_t u = u0; u < u1; u++) cnt2[(size_t)u * PART_B + b] += bst;
    if (b < nsub) nbound[((size_t)s << step) + b] = bst;
    if (s == nseg - 1 && b == nsub - 1) nbound[(size_t)nseg << step] = bst + run;
}
__global__ void __launch_bounds__(PART_BLOCK)
    k_seg_move(GridC g, SpeciesBuf src, SpeciesBuf dst, const uint32_t* __restrict__ bound,
               const uint32_t* __restrict__ tbase, int nseg, int shift, int step, const uint32_t* __restrict__ cnt2) {
    int s, cnt;
    size_t base;
    if (!seg_tile(bound, tbase, nseg, blockIdx.x, s, base, cnt)) return;  // (whole workgroup)
    part_tile(g, src, dst, base, cnt, shift, (uint32_t)s << step, 1 << step, cnt2 + (size_t)blockIdx.x * PART_B, 1);
}

// final buckets: bucket f = [bound[f], bound[f + 1]) holds the keys
// [f << shift, (f + 1) << shift), at most FINE_CELLS of them.  Writes off[]
// for those cells (and the live count with the removed particles' key).
__global__ void __launch_bounds__(PART_BLOCK)
    k_fine_sort(GridC g, SpeciesBuf src, SpeciesBuf dst, const uint32_t* __restrict__ bound, int shift,
                uint32_t* __restrict__ off, uint32_t* __restrict__ gcount, uint32_t* __restrict__ grank,
                uint32_t* __restrict__ d_n) {
    __shared__ uint32_t lcnt[FINE_CELLS];
    __shared__ uint16_t lpos[FINE_CAP];
    __shared__ uint32_t stage[FINE_CAP];
    const uint32_t f = blockIdx.x, fs = bound[f], fe = bound[f + 1];
    const uint32_t k0 = f << shift;
    const uint32_t k1 = min((f + 1) << shift, (uint32_t)g.ncells + 1);  // keys of this bucket
    if (k0 >= k1) return;  // (whole workgroup)
    const int ncl = (int)(k1 - k0), size = (int)(fe - fs);
    const bool lds = size <= FINE_CAP;
    for (int k = threadIdx.x; k < ncl; k += PART_BLOCK) lcnt[k] = 0;
    __syncthreads();
    if (lds) {
        for (int q = threadIdx.x; q < size; q += PART_BLOCK) {
            const uint32_t k = cell_key(g, src.cell[fs + q]) - k0;
            lpos[q] = (uint16_t)atomicAdd(&lcnt[k], 1u);
        }
    } else {  // ranks from global atomics (gcount zeroed), kept in grank
        for (int q = threadIdx.x; q < size; q += PART_BLOCK) {
            const uint32_t k = cell_key(g, src.cell[fs + q]) - k0;
            grank[fs + q] = atomicAdd(&gcount[k0 + k], 1u);
            atomicAdd(&lcnt[k], 1u);
        }
    }
    __syncthreads();
    {  // exclusive scan of the keys' counts, FINE_CELLS / PART_BLOCK per thread
        constexpr int PER = FINE_CELLS / PART_BLOCK;
        uint32_t v[PER], sum = 0;
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int k = threadIdx.x * PER + e;
            v[e] = k < ncl ? lcnt[k] : 0u;
            sum += v[e];
        }
        uint32_t ex = block_exclusive_scan(sum, nullptr);
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int k = threadIdx.x * PER + e;
            if (k < ncl) {
                lcnt[k] = ex;
                const uint32_t key = k0 + (uint32_t)k;
                off[key] = fs + ex;
                if (key == (uint32_t)g.ncells) {  // the removed particles' bucket: the live count
                    off[key + 1] = fe;
                    d_n[0] = fs + ex;
                    d_n[3] = fs + ex;
                }
            }
            ex += v[e];
        }
    }
    __syncthreads();
    const uint32_t* in[8] = {src.cell,
                             reinterpret_cast<const uint32_t*>(src.f[0]),
                             reinterpret_cast<const uint32_t*>(src.f[1]),
                             reinterpret_cast<const uint32_t*>(src.f[2]),
                             reinterpret_cast<const uint32_t*>(src.v[0]),
                             reinterpret_cast<const uint32_t*>(src.v[1]),
                             reinterpret_cast<const uint32_t*>(src.v[2]),
                             reinterpret_cast<const uint32_t*>(src.w)};
    uint32_t* out[8] = {dst.cell,
                        reinterpret_cast<uint32_t*>(dst.f[0]),
                        reinterpret_cast<uint32_t*>(dst.f[1]),
                        reinterpret_cast<uint32_t*>(dst.f[2]),
                        reinterpret_cast<uint32_t*>(dst.v[0]),
                        reinterpret_cast<uint32_t*>(dst.v[1]),
                        reinterpret_cast<uint32_t*>(dst.v[2]),
                        reinterpret_cast<uint32_t*>(dst.w)};
    if (lds) {
        // each thread's PER slots (q = threadIdx.x + e * PART_BLOCK): final
        // positions in registers, every array's loads issued before their
        // LDS stores
        constexpr int PER = FINE_CAP / PART_BLOCK;
        uint32_t mypos[PER];
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const int q = threadIdx.x + e * PART_BLOCK;
            mypos[e] = q < size ? lcnt[cell_key(g, src.cell[fs + q]) - k0] + lpos[q] : 0u;
        }
#pragma unroll 1
        for (int a = 0; a < 8; a++) {
            uint32_t v[PER];
#pragma unroll
            for (int e = 0; e < PER; e++) {
                const int q = threadIdx.x + e * PART_BLOCK;
                v[e] = q < size ? in[a][fs + q] : 0u;
            }
#pragma unroll
            for (int e = 0; e < PER; e++)
                if (threadIdx.x + e * PART_BLOCK < size) stage[mypos[e]] = v[e];
            __syncthreads();
#pragma unroll
            for (int e = 0; e < PER; e++) {
                const int r = threadIdx.x + e * PART_BLOCK;
                if (r < size) out[a][fs + r] = stage[r];
            }
            __syncthreads();
        }
    } else {
        for (int q = threadIdx.x; q < size; q += PART_BLOCK) {
            const uint32_t k = cell_key(g, src.cell[fs + q]) - k0;
            const uint32_t d = fs + lcnt[k] + grank[fs + q];
#pragma unroll
            for (int a = 0; a < 8; a++) out[a][d] = in[a][fs + q];
        }
    }
}

// Second pass of pic_push_sort: every particle is pushed again from the
// unchanged source arrays (the same arithmetic as the first pass, so the
// same new cell) and written straight to its sorted position off[new cell] +
// its slot in the cell.  The slot comes from the first pass's run bases of
// the particle's group of 64 (same groups, same keys, same runs): no atomics.
// Removed particles (key ncells) are dropped.
// ILP particles per thread, BLOCK apart (all loads issued before any wait);
// the groups of 64 consecutive particles are the first pass's.  One per
// thread: two measured 10% slower (0.318 against 0.289 ms for 20M)
constexpr int PUSH_SCATTER_ILP = 1;
template <bool SURF, int ILP>
__global__ void __launch_bounds__(BLOCK)
    k_push_scatter(GridC g, SpeciesBuf src, SpeciesBuf dst, const float4* __restrict__ E, PushArgs a,
                   const uint32_t* __restrict__ bases, const uint32_t* __restrict__ off,
                   const uint32_t* __restrict__ d_n, uint32_t cap) {
    const size_t p0 = xcd_wg() * (BLOCK * ILP) + threadIdx.x;
    const uint32_t n = d_n[1];
    uint32_t c[ILP];
    float f[ILP][3], v[ILP][3], w[ILP];
#pragma unroll
    for (int q = 0; q < ILP; q++) {  // every load before any wait
        const size_t p = p0 + (size_t)q * BLOCK;
        c[q] = DEAD;
        w[q] = 0.f;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            f[q][d] = 0.f;
            v[q][d] = 0.f;
        }
        if (p < n) {
            c[q] = src.cell[p];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                f[q][d] = src.f[d][p];
                v[q][d] = src.v[d][p];
            }
            w[q] = src.w[p];
        }
    }
#pragma unroll
    for (int q = 0; q < ILP; q++) {
        const size_t p = p0 + (size_t)q * BLOCK;
        const bool valid = p < n;
        uint32_t nc = DEAD;
        if (c[q] != DEAD) nc = advance<SURF>(g, E, a, c[q], f[q], v[q], src, p);
        // the first pass's key: removed particles in bucket ncells, lanes past n none
        const uint32_t key = valid ? cell_key(g, nc) : 0xFFFFFFFFu;
        int run, head_lane, len;
        bool head;
        key_runs(key, run, head_lane, len, head);  // every lane of the wave takes part
        if (!valid || nc == DEAD) continue;
        const uint32_t slot =
            bases[(p & ~(size_t)63) + run] + (uint32_t)((int)(threadIdx.x & 63) - head_lane);
        const uint32_t dst_i = off[key] + slot;
        if (dst_i >= cap) continue;  // cannot happen with a current histogram
        dst.cell[dst_i] = nc;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            dst.f[d][dst_i] = f[q][d];
            dst.v[d][dst_i] = v[q][d];
        }
        dst.w[dst_i] = w[q];
    }
}

// Number density without atomics, in two passes over the sorted particles.
// Pass 1, one thread per cell: the cell's particles (contiguous after the
// sort, so the cell word is not read) are folded into the eight corner sums
// of the CIC weights w * W_q, stored per corner as corner[q][cell] (FP32).
// Pass 2, one thread per node: den = (1 / V_node) * the sum, over the up to
// eight cells around the node, of that cell's sum for this corner (plain
// store; Species.cpp:401-416 scatters the same weights with atomics).
// Corner q = (dk*2 + dj)*2 + di of cell (i, j, k) is node (i+di, j+dj, k+dk).
// Pass 1 stages a workgroup's particles through LDS: the workgroup owns
// BLOCK consecutive cells, whose particles form one contiguous range; it
// loads the range in chunks of DEP_CHUNK particles with 16-byte loads (all
// issued before the first LDS store), then every thread folds the part of
// its cell's particles that lies in the chunk, in particle order.
constexpr int DEP_CHUNK = 2048;
__global__ void __launch_bounds__(BLOCK)
    k_dep_cells(GridC g, SpeciesBuf b, const uint32_t* __restrict__ off, float* __restrict__ corner) {
    __shared__ float4 sf[4][DEP_CHUNK / 4];  // fx, fy, fz, w of the chunk
    const int tid = threadIdx.x;
    const int c = blockIdx.x * BLOCK + tid;
    const int clast = min((int)(blockIdx.x + 1) * BLOCK, g.ncells);  // one past the block's cells
    const uint32_t P0 = off[blockIdx.x * BLOCK], P1 = off[clast];
    uint32_t p0 = 0, p1 = 0;
    if (c < g.ncells) {
        p0 = off[c];
        p1 = off[c + 1];
    }
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    // chunk starts are multiples of 4 (16-byte loads); the particle arrays
    // are padded to 64 entries, so a group of 4 never leaves its array
    for (uint32_t cs = P0 & ~3u; cs < P1; cs += DEP_CHUNK) {
        constexpr int PER = DEP_CHUNK / 4 / BLOCK;  // float4 per array per thread
        float4 r[4][PER];
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const uint32_t q = cs + 4 * (tid + BLOCK * e);
            const bool in = q < P1;
            r[0][e] = in ? *reinterpret_cast<const float4*>(b.f[0] + q) : make_float4(0.f, 0.f, 0.f, 0.f);
            r[1][e] = in ? *reinterpret_cast<const float4*>(b.f[1] + q) : make_float4(0.f, 0.f, 0.f, 0.f);
            r[2][e] = in ? *reinterpret_cast<const float4*>(b.f[2] + q) : make_float4(0.f, 0.f, 0.f, 0.f);
            r[3][e] = in ? *reinterpret_cast<const float4*>(b.w + q) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __syncthreads();  // the previous chunk is consumed
#pragma unroll
        for (int e = 0; e < PER; e++)
#pragma unroll
            for (int a = 0; a < 4; a++) sf[a][tid + BLOCK * e] = r[a][e];
        __syncthreads();
        const float* fx_ = reinterpret_cast<const float*>(sf[0]);
        const float* fy_ = reinterpret_cast<const float*>(sf[1]);
        const float* fz_ = reinterpret_cast<const float*>(sf[2]);
        const float* w_ = reinterpret_cast<const float*>(sf[3]);
        const uint32_t ce = cs + DEP_CHUNK;
        const uint32_t lo = p0 > cs ? p0 : cs, hi = p1 < ce ? p1 : ce;
        for (uint32_t p = lo; p < hi; p++) {
            const uint32_t l = p - cs;
            float fx = fx_[l], fy = fy_[l], fz = fz_[l], w = w_[l];
            float gx = 1.f - fx, gy = 1.f - fy, gz = 1.f - fz;
            float wgz = w * gz, wfz = w * fz;
            acc[0] += gx * gy * wgz;
            acc[1] += fx * gy * wgz;
            acc[2] += gx * fy * wgz;
            acc[3] += fx * fy * wgz;
            acc[4] += gx * gy * wfz;
            acc[5] += fx * gy * wfz;
            acc[6] += gx * fy * wfz;
            acc[7] += fx * fy * wfz;
        }
    }
    if (c >= g.ncells) return;
#pragma unroll
    for (int q = 0; q < 8; q++) corner[(size_t)q * g.ncells + c] = acc[q];
}

__global__ void __launch_bounds__(BLOCK)
    k_dep_nodes(GridC g, const float* __restrict__ corner, double* __restrict__ den) {
    long long u = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (u >= g.nn) return;
    unsigned uu = (unsigned)u, row = uu / (unsigned)g.ni, kk = row / (unsigned)g.nj;
    int i = (int)(uu - row * (unsigned)g.ni), j = (int)(row - kk * (unsigned)g.nj), k = (int)kk;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        int ci = i - (q & 1), cj = j - ((q >> 1) & 1), ck = k - (q >> 2);
        if (ci < 0 || ci >= g.ci || cj < 0 || cj >= g.cj || ck < 0 || ck >= g.ck) continue;
        s += (double)corner[(size_t)q * g.ncells + cell_index(g, ci, cj, ck)];
    }
    den[u] = s * inv_node_vol(g, i, j, k);
}

// Velocity moments (Species.cpp:767-776: n, nv, nuu, nvv, nww accumulated
// with CIC weights) in the same two passes as the density: per cell, the
// seven channels' eight corner sums (FP32, particle order) from LDS-staged
// chunks; per node, the sum over its cells added to the FP64 moments (the
// node's thread owns it: no atomics), and with want_den the density.
constexpr int MOM_CHUNK = 1024;
constexpr int MOM_CH = 7;
__global__ void __launch_bounds__(BLOCK)
    k_mom_cells(GridC g, SpeciesBuf b, const uint32_t* __restrict__ off, float* __restrict__ corner) {
    __shared__ float4 sf[7][MOM_CHUNK / 4];  // fx, fy, fz, w, vx, vy, vz of the chunk
    const int tid = threadIdx.x;
    const int c = blockIdx.x * BLOCK + tid;
    const int clast = min((int)(blockIdx.x + 1) * BLOCK, g.ncells);
    const uint32_t P0 = off[blockIdx.x * BLOCK], P1 = off[clast];
    uint32_t p0 = 0, p1 = 0;
    if (c < g.ncells) {
        p0 = off[c];
        p1 = off[c + 1];
    }
    float acc[MOM_CH][8];
#pragma unroll
    for (int ch = 0; ch < MOM_CH; ch++)
#pragma unroll
        for (int q = 0; q < 8; q++) acc[ch][q] = 0.f;
    const float* src[7] = {b.f[0], b.f[1], b.f[2], b.w, b.v[0], b.v[1], b.v[2]};
    for (uint32_t cs = P0 & ~3u; cs < P1; cs += MOM_CHUNK) {
        constexpr int PER = MOM_CHUNK / 4 / BLOCK;
        float4 r[7][PER];
#pragma unroll
        for (int e = 0; e < PER; e++) {
            const uint32_t q = cs + 4 * (tid + BLOCK * e);
            const bool in = q < P1;  // arrays padded to 64 entries: a quad never leaves its array
#pragma unroll
            for (int a = 0; a < 7; a++)
                r[a][e] = in ? *reinterpret_cast<const float4*>(src[a] + q) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __syncthreads();  // the previous chunk is consumed
#pragma unroll
        for (int e = 0; e < PER; e++)
#pragma unroll
            for (int a = 0; a < 7; a++) sf[a][tid + BLOCK * e] = r[a][e];
        __syncthreads();
        const float* sv[7];
#pragma unroll
        for (int a = 0; a < 7; a++) sv[a] = reinterpret_cast<const float*>(sf[a]);
        const uint32_t ce = cs + MOM_CHUNK;
        const uint32_t lo = p0 > cs ? p0 : cs, hi = p1 < ce ? p1 : ce;
        for (uint32_t p = lo; p < hi; p++) {
            const uint32_t l = p - cs;
            const float fx = sv[0][l], fy = sv[1][l], fz = sv[2][l], w = sv[3][l];
            const float vx = sv[4][l], vy = sv[5][l], vz = sv[6][l];
            const float gx = 1.f - fx, gy = 1.f - fy, gz = 1.f - fz;
            float wt[8];  // corner (dk*2 + dj)*2 + di
            wt[0] = gx * gy * gz;
            wt[1] = fx * gy * gz;
            wt[2] = gx * fy * gz;
            wt[3] = fx * fy * gz;
            wt[4] = gx * gy * fz;
            wt[5] = fx * gy * fz;
            wt[6] = gx * fy * fz;
            wt[7] = fx * fy * fz;
            const float val[MOM_CH] = {w, w * vx, w * vy, w * vz, w * vx * vx, w * vy * vy, w * vz * vz};
#pragma unroll
            for (int ch = 0; ch < MOM_CH; ch++)
#pragma unroll
                for (int q = 0; q < 8; q++) acc[ch][q] += val[ch] * wt[q];
        }
    }
    if (c >= g.ncells) return;
#pragma unroll
    for (int ch = 0; ch < MOM_CH; ch++)
#pragma unroll
        for (int q = 0; q < 8; q++) corner[((size_t)ch * 8 + q) * g.ncells + c] = acc[ch][q];
}

__global__ void __launch_bounds__(BLOCK)
    k_mom_nodes(GridC g, const float* __restrict__ corner, double* __restrict__ mom, double* __restrict__ den) {
    long long u = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (u >= g.nn) return;
    unsigned uu = (unsigned)u, row = uu / (unsigned)g.ni, kk = row / (unsigned)g.nj;
    int i = (int)(uu - row * (unsigned)g.ni), j = (int)(row - kk * (unsigned)g.nj), k = (int)kk;
    long long cell[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        int ci = i - (q & 1), cj = j - ((q >> 1) & 1), ck = k - (q >> 2);
        cell[q] = (ci < 0 || ci >= g.ci || cj < 0 || cj >= g.cj || ck < 0 || ck >= g.ck)
                      ? -1
                      : (long long)cell_index(g, ci, cj, ck);
    }
#pragma unroll
    for (int ch = 0; ch < MOM_CH; ch++) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (cell[q] >= 0) s += (double)corner[((size_t)ch * 8 + q) * g.ncells + cell[q]];
        mom[(long long)ch * g.nn + u] += s;
        if (ch == 0 && den) den[u] = s * inv_node_vol(g, i, j, k);
    }
}

// The appended part of a two-part species as an index: tidx[toff[cell] +
// rank] = the particle's position, for [start, d_n[1]) (DEAD ones left out).
__global__ void __launch_bounds__(BLOCK)
    k_tail_index(GridC g, const uint32_t* __restrict__ cell, const uint32_t* __restrict__ rank,
                 const uint32_t* __restrict__ toff, const uint32_t* __restrict__ d_n, uint32_t start,
                 uint32_t* __restrict__ tidx) {
    const size_t p = (size_t)start + (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= d_n[1]) return;
    const uint32_t c = cell[p];
    if (c == DEAD) return;
    tidx[toff[cell_key(g, c)] + rank[p]] = (uint32_t)p;
}

// The index extended by the particles appended since it was built (their
// per-cell offsets boff, ranks from k_hist): a cell's previous entries move
// by boff[c] (toff_n[c] = toff[c] + boff[c]); the new particles follow them
// (k_tail_index_more).  A workgroup takes BLOCK consecutive cells, whose
// previous entries are one contiguous range: the threads stream it and find
// each entry's cell by binary search over the cells' offsets in LDS.
__global__ void __launch_bounds__(BLOCK)
    k_tail_merge(const uint32_t* __restrict__ toff, const uint32_t* __restrict__ boff, int ncells,
                 const uint32_t* __restrict__ tidx, uint32_t* __restrict__ toff_n, uint32_t* __restrict__ tidx_n) {
    __shared__ uint32_t so[BLOCK + 1], sb[BLOCK];
    const size_t c0 = (size_t)xcd_wg() * BLOCK, c = c0 + threadIdx.x;
    if (c <= (size_t)ncells + 1) {
        const uint32_t o = toff[c], b = boff[c];
        toff_n[c] = o + b;
        so[threadIdx.x] = o;
        sb[threadIdx.x] = b;
    }
    if (c0 >= (size_t)ncells) return;  // (whole workgroup)
    const int nc = (int)min((size_t)BLOCK, (size_t)ncells - c0);  // cells whose entries move
    if (threadIdx.x == 0) so[nc] = toff[c0 + nc];
    __syncthreads();
    const uint32_t e1 = so[nc];
    for (uint32_t j = so[0] + threadIdx.x; j < e1; j += BLOCK) {
        int lo = 0, hi = nc - 1;  // the last cell k with so[k] <= j
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (so[mid] <= j)
                lo = mid;
            else
                hi = mid - 1;
        }
        tidx_n[j + sb[lo]] = tidx[j];
    }
}
__global__ void __launch_bounds__(BLOCK)
    k_tail_index_more(GridC g, const uint32_t* __restrict__ cell, const uint32_t* __restrict__ rank,
                      const uint32_t* __restrict__ toff, const uint32_t* __restrict__ boff,
                      const uint32_t* __restrict__ d_n, uint32_t start, uint32_t* __restrict__ tidx_n) {
    const size_t p = (size_t)start + (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= d_n[1]) return;
    const uint32_t c = cell[p];
    if (c == DEAD) return;
    const uint32_t key = cell_key(g, c);
    tidx_n[toff[key + 1] + boff[key] + rank[p]] = (uint32_t)p;
}

// The two parts joined into one cell-sorted array: a cell's particles of the
// sorted part, then its appended ones in index order (the order of a cell's
// pick pool).  The sorted part's particle i moves to i + toff[cell]; the
// appended part's j-th indexed particle to off[cell + 1] + j.  Both reads
// and writes stream (the appended part is gathered through the index).
__device__ inline void copy_particle(const SpeciesBuf& src, size_t p, const SpeciesBuf& dst, size_t q,
                                     uint32_t c) {
    dst.cell[q] = c;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        dst.f[d][q] = src.f[d][p];
        dst.v[d][q] = src.v[d][p];
    }
    dst.w[q] = src.w[p];
}
__global__ void __launch_bounds__(BLOCK)
    k_join_sorted(GridC g, SpeciesBuf src, SpeciesBuf dst, const uint32_t* __restrict__ toff, uint32_t nb) {
    const size_t i = (size_t)xcd_wg() * BLOCK + threadIdx.x;
    if (i >= nb) return;
    const uint32_t c = src.cell[i];
    if (c == DEAD) return;  // none in a sorted part
    copy_particle(src, i, dst, i + toff[cell_key(g, c)], c);
}
__global__ void __launch_bounds__(BLOCK)
    k_join_tail(GridC g, SpeciesBuf src, SpeciesBuf dst, const uint32_t* __restrict__ off,
                const uint32_t* __restrict__ toff, const uint32_t* __restrict__ tidx) {
    const size_t j = (size_t)xcd_wg() * BLOCK + threadIdx.x;
    if (j >= toff[g.ncells]) return;
    const uint32_t p = tidx[j];
    const uint32_t c = src.cell[p];
    copy_particle(src, p, dst, off[cell_key(g, c) + 1] + j, c);
}
// the joined offsets (off[c] + toff[c]) and live count
__global__ void __launch_bounds__(BLOCK)
    k_join_off(uint32_t* __restrict__ off, const uint32_t* __restrict__ toff, int ncells,
               uint32_t* __restrict__ d_n) {
    const size_t c = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (c > (size_t)ncells) return;
    const uint32_t v = off[c] + toff[c];
    off[c] = v;
    if (c == (size_t)ncells) {
        off[ncells + 1] = v;
        d_n[0] = v;
        d_n[3] = v;
    }
}

struct RhoArgs {
    const double* den[MAX_SPECIES];
    double q[MAX_SPECIES];
    int ns;
};

__global__ void __launch_bounds__(BLOCK) k_rho(long long nn, RhoArgs a, double* __restrict__ rho) {
    long long u = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (u >= nn) return;
    double r = 0;
    for (int s = 0; s < a.ns; s++) r += a.q[s] * a.den[s][u];
    rho[u] = r;
}

struct LoadArgs {
    double xmin[3], xmax[3];
    double vth;     // sqrt(2kT/m) (mode 0) or sqrt(kT/m) (mode 1)
    double rewind;  // (q/m) * 0.5 dt
    uint32_t seed, seq;
    int mode;
    float w;
};

__global__ void __launch_bounds__(BLOCK)
    k_load(GridC g, SpeciesBuf b, const float4* __restrict__ E, LoadArgs a, size_t base,
           size_t count) {
    size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= count) return;
    Draws dr;
    dr.init((uint32_t)i, (uint32_t)(i >> 32), a.seed, a.seq, TAG_LOAD);
    double pos[3], vel[3];
    for (int d = 0; d < 3; d++) pos[d] = a.xmin[d] + dr(d) * (a.xmax[d] - a.xmin[d]);
    if (a.mode == 0) {  // Species.cpp:855-869
        double v[3];
        for (int c = 0; c < 3; c++) {
            double u1 = dr(3 + 3 * c), u2 = dr(4 + 3 * c), u3 = dr(5 + 3 * c);
            v[c] = a.vth * (u1 + u2 + u3 - 1.5);
        }
        double sp = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        double theta = 2 * PI_REF * dr(12);
        double r = -1.0 + 2 * dr(13);
        double s = sqrt(1 - r * r);
        vel[0] = sp * r;
        vel[1] = sp * (cos(theta) * s);
        vel[2] = sp * (sin(theta) * s);
    } else {  // Box-Muller Maxwellian
        double u1 = dr(3), u2 = dr(4), u3 = dr(5), u4 = dr(6);
        double r1 = sqrt(-2.0 * log(1.0 - u1)), r2 = sqrt(-2.0 * log(1.0 - u3));
        vel[0] = a.vth * r1 * cos(TWO_PI * u2);
        vel[1] = a.vth * r1 * sin(TWO_PI * u2);
        vel[2] = a.vth * r2 * cos(TWO_PI * u4);
    }
    // addParticle (Species.cpp:420-434): reject, then rewind v by half a step
    double lc[3];
    int ijk[3];
    int lim[3] = {g.ci, g.cj, g.ck};
    float f[3] = {0, 0, 0};
    uint32_t c = 0;
    for (int d = 0; d < 3; d++) {
        lc[d] = (pos[d] - g.x0[d]) * g.inv_dx[d];
        if (!(lc[d] >= 0.0)) c = DEAD;
        ijk[d] = (int)lc[d];
        if (ijk[d] >= lim[d]) c = DEAD;
        float fr = (float)(lc[d] - ijk[d]);
        f[d] = fr < 1.0f ? fr : 0.99999994f;
    }
    if (c != DEAD && g.nrect && in_rect(g, ijk[0], ijk[1], ijk[2], f[0], f[1], f[2])) c = DEAD;
    if (c != DEAD) {
        c = pack_cell(ijk[0], ijk[1], ijk[2]);
        if (a.rewind != 0.0) {
            float3 e = gather_e(g, E, ijk[0], ijk[1], ijk[2], f[0], f[1], f[2]);
            vel[0] -= (double)e.x * a.rewind;
            vel[1] -= (double)e.y * a.rewind;
            vel[2] -= (double)e.z * a.rewind;
        }
    }
    size_t o = base + i;
    b.cell[o] = c;
    for (int d = 0; d < 3; d++) {
        b.f[d][o] = f[d];
        b.v[d][o] = (float)vel[d];
    }
    b.w[o] = a.w;
}

// Variable-weight merge (Species::merge, Species.cpp:987-1145) on the
// cell-sorted arrays, one workgroup per cell of at least 10 particles
// (:1029): the cell's velocity range widened by 0.1% at the maximum
// (:1043-1057) is split into 15 bins per axis (Species.h:129, :1059-1067);
// the workgroup sorts (bin, place in the cell) keys in LDS (bitonic), and
// one thread per bin of more than two particles sums weight, momentum and
// per-axis energy in cell order in FP64, and writes the pair
// v = mean +- sqrt(variance) (a negative variance clamped to 0), half the
// weight each, at the positions of two distinct random members, in place
// of the bin's first two members; the other members are removed (DEAD).
// Charged species get addParticle's half-step rewind (:1141-1142,
// Species.cpp:427-433).  Draws: Philox (seed, call), counter (cell, bin).
constexpr int MERGE_BLOCK = 256;
constexpr int MERGE_MAX = 8192;  // particles of one cell (LDS keys)
constexpr int MERGE_BINS = 15;
constexpr int MERGE_IDX_BITS = 13;
struct MergeArgs {
    uint32_t seed, call;
    double rewind;  // (q/m) * 0.5 dt, 0 for neutrals
};
__global__ void __launch_bounds__(MERGE_BLOCK)
    k_merge(GridC g, SpeciesBuf b, const uint32_t* __restrict__ off, const float4* __restrict__ E,
            MergeArgs a, unsigned long long* __restrict__ stats) {
    __shared__ uint32_t key[MERGE_MAX];
    __shared__ double red[6][MERGE_BLOCK / 64];
    const int c = blockIdx.x;
    const int tid = threadIdx.x;
    const uint32_t p0 = off[c], p1 = off[c + 1];
    const int n = (int)(p1 - p0);
    if (n < 10) return;
    if (n > MERGE_MAX) {
        if (tid == 0) atomicAdd(&stats[2], 1ull);
        return;
    }
    // velocity range of the cell (exact min and max)
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int q = tid; q < n; q += MERGE_BLOCK)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double v = (double)b.v[d][p0 + q];
            mn[d] = fmin(mn[d], v);
            mx[d] = fmax(mx[d], v);
        }
#pragma unroll
    for (int d = 0; d < 3; d++)
        for (int o = 32; o > 0; o >>= 1) {
            mn[d] = fmin(mn[d], __shfl_xor(mn[d], o, 64));
            mx[d] = fmax(mx[d], __shfl_xor(mx[d], o, 64));
        }
    if ((tid & 63) == 0)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            red[d][tid >> 6] = mn[d];
            red[3 + d][tid >> 6] = mx[d];
        }
    __syncthreads();
    double vmin[3], dv[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double lo = red[d][0], hi = red[3 + d][0];
        for (int w = 1; w < MERGE_BLOCK / 64; w++) {
            lo = fmin(lo, red[d][w]);
            hi = fmax(hi, red[3 + d][w]);
        }
        if (hi > 0)
            hi *= 1.001;
        else
            hi *= 0.999;
        vmin[d] = lo;
        dv[d] = (hi - lo) / (double)MERGE_BINS;
    }
    // keys: bin << 13 | place; padded to a power of two
    int npad = 16;
    while (npad < n) npad <<= 1;
    for (int q = tid; q < npad; q += MERGE_BLOCK) {
        uint32_t k = 0xFFFFFFFFu;
        if (q < n) {
            int bi[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double x = ((double)b.v[d][p0 + q] - vmin[d]) / dv[d];
                int t = x == x ? (x < (double)MERGE_BINS ? (int)x : MERGE_BINS - 1) : 0;
                bi[d] = t < 0 ? 0 : t;
            }
            const uint32_t bin = (uint32_t)((bi[0] * MERGE_BINS + bi[1]) * MERGE_BINS + bi[2]);
            k = (bin << MERGE_IDX_BITS) | (uint32_t)q;
        }
        key[q] = k;
    }
    __syncthreads();
    for (int k = 2; k <= npad; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < npad; i += MERGE_BLOCK) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint32_t x = key[i], y = key[ixj];
                    if ((x > y) == ((i & k) == 0)) {
                        key[i] = y;
                        key[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    constexpr uint32_t MASK = (1u << MERGE_IDX_BITS) - 1u;
    unsigned long long nbins = 0, nin = 0;
    for (int s0 = tid; s0 < n; s0 += MERGE_BLOCK) {
        const uint32_t bin = key[s0] >> MERGE_IDX_BITS;
        if (s0 > 0 && (key[s0 - 1] >> MERGE_IDX_BITS) == bin) continue;  // not a bin's first
        int e = s0 + 1;
        while (e < n && (key[e] >> MERGE_IDX_BITS) == bin) e++;
        const int len = e - s0;
        if (len <= 2) continue;
        nbins++;
        nin += (unsigned long long)len;
        double tw = 0, tm[3] = {0, 0, 0}, te[3] = {0, 0, 0};
        for (int q = s0; q < e; q++) {
            const uint32_t idx = p0 + (key[q] & MASK);
            const double w = (double)b.w[idx];
            tw += w;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double v = (double)b.v[d][idx];
                tm[d] += v * w;
                te[d] += (v * v) * w;
            }
        }
        const double wh = tw * 0.5, inv = 1.0 / tw;
        double va[3], vb[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            tm[d] *= inv;
            te[d] *= inv;
            te[d] -= tm[d] * tm[d];
            if (te[d] < 0) te[d] = 0;
            const double sq = sqrt(te[d]);
            va[d] = tm[d] + sq;
            vb[d] = tm[d] - sq;
        }
        Draws dr;
        dr.init((uint32_t)c, bin, a.seed, a.call, TAG_MERGE);
        int slot = 0;
        const int ia = (int)(dr(slot++) * (double)len);
        int ib = (int)(dr(slot++) * (double)len);
        while (ib == ia) ib = (int)(dr(slot++) * (double)len);
        const uint32_t xa = p0 + (key[s0 + ia] & MASK), xb = p0 + (key[s0 + ib] & MASK);
        const uint32_t ca = b.cell[xa], cb = b.cell[xb];
        float fa[3], fb[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            fa[d] = b.f[d][xa];
            fb[d] = b.f[d][xb];
        }
        if (a.rewind != 0.0) {
            int i, j, k;
            unpack_cell(ca, i, j, k);
            const float3 ea = gather_e(g, E, i, j, k, fa[0], fa[1], fa[2]);
            unpack_cell(cb, i, j, k);
            const float3 eb = gather_e(g, E, i, j, k, fb[0], fb[1], fb[2]);
            va[0] -= (double)ea.x * a.rewind;
            va[1] -= (double)ea.y * a.rewind;
            va[2] -= (double)ea.z * a.rewind;
            vb[0] -= (double)eb.x * a.rewind;
            vb[1] -= (double)eb.y * a.rewind;
            vb[2] -= (double)eb.z * a.rewind;
        }
        const uint32_t da = p0 + (key[s0] & MASK), db = p0 + (key[s0 + 1] & MASK);
        b.cell[da] = ca;
        b.cell[db] = cb;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            b.f[d][da] = fa[d];
            b.f[d][db] = fb[d];
            b.v[d][da] = (float)va[d];
            b.v[d][db] = (float)vb[d];
        }
        b.w[da] = (float)wh;
        b.w[db] = (float)wh;
        for (int q = s0 + 2; q < e; q++) {
            const uint32_t idx = p0 + (key[q] & MASK);
            b.cell[idx] = DEAD;
            b.w[idx] = 0.0f;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        nbins += __shfl_xor(nbins, o, 64);
        nin += __shfl_xor(nin, o, 64);
    }
    if ((tid & 63) == 0 && nbins) {
        atomicAdd(&stats[0], nbins);
        atomicAdd(&stats[1], nin);
    }
}

}  // namespace

// ------------------------------------------------------------------------
// host entry points
// ------------------------------------------------------------------------
static PushArgs push_args(pic_ctx* ctx, Species& s, double dt) {
    PushArgs a;
    a.qmdt = (float)(dt * s.charge / s.mass);
    for (int d = 0; d < 3; d++) a.dtx[d] = (float)(dt * ctx->g.inv_dx[d]);
    a.kick = s.charge != 0.0;
    // without objects a particle that stays in its cell cannot be removed
    a.emax = ctx->g.nrect ? nullptr : ctx->emax;
    SurfArgs& f = a.s;
    f = SurfArgs{};
    f.mode = ctx->g.nrect ? s.surf_mode : PIC_SURF_ABSORB;
    f.dt = dt;
    f.vth_wall = sqrt(2 * KB * 300.0 / s.mass);  // Species.cpp:836, :856
    f.seed = s.surf_seed;
    f.step = s.surf_step++;
    if (f.mode == PIC_SURF_RECYCLE) {
        Species& t = ctx->sp[s.surf_target];
        f.mpw0_t = t.mpw0;
        f.tgt = t.buf[t.cur];
        f.tgt_n = t.d_n;
        f.tgt_cap = (uint32_t)t.cap;
    }
    return a;
}

// after a push of species s that may have emitted particles into its
// target: the target's live count (its launch sizes), its flags (only
// appended: a sorted target keeps its sorted part and may sort the tail
// alone), and the emissions that did not fit
static int after_emission(pic_ctx* ctx, const Species& s, const PushArgs& a) {
    if (a.s.mode != PIC_SURF_RECYCLE) return PIC_OK;
    Species& t = ctx->sp[s.surf_target];
    uint32_t hn[4];
    PIC_HIP(ctx, hipMemcpyAsync(hn, t.d_n, sizeof(hn), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (hn[0] > t.cap) hn[0] = (uint32_t)t.cap;
    t.bulk_n = hn[3];  // the emissions were appended behind the last whole sort's particles
    if ((size_t)hn[0] != t.n_upper) {
        t.n_upper = hn[0];
        t.sorted = false;
        t.hist_valid = false;
    }
    if (hn[2]) {
        uint32_t z = 0;
        PIC_HIP(ctx, hipMemcpyAsync(t.d_n + 2, &z, sizeof(z), hipMemcpyHostToDevice, ctx->stream));
        PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        PIC_HIP(ctx, hipMemcpyAsync(t.d_n, &hn[0], sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return pic_fail(ctx, PIC_ERR_CAPACITY, "particles emitted at objects exceed the target's capacity");
    }
    return PIC_OK;
}

extern "C" int pic_species_set_surface(pic_ctx* ctx, int sid, int mode, int sid_target, uint32_t seed) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    if (mode != PIC_SURF_ABSORB && mode != PIC_SURF_REFLECT && mode != PIC_SURF_RECYCLE)
        return pic_fail(ctx, PIC_ERR_ARG, "unknown surface mode");
    if (mode == PIC_SURF_RECYCLE) {
        rc = pic_check_sid(ctx, sid_target);
        if (rc) return rc;
        if (sid_target == sid) return pic_fail(ctx, PIC_ERR_ARG, "a species cannot emit into itself");
    }
    Species& s = ctx->sp[sid];
    s.surf_mode = mode;
    s.surf_target = mode == PIC_SURF_RECYCLE ? sid_target : -1;
    s.surf_seed = seed;
    return PIC_OK;
}

extern "C" int pic_push(pic_ctx* ctx, int sid, double dt) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    Species& s = ctx->sp[sid];
    const GridC& g = ctx->g;
    PushArgs a = push_args(ctx, s, dt);
    a.s.emit = 1;
    PIC_HIP(ctx, hipMemsetAsync(s.count, 0, ((size_t)g.ncells + 1) * 4, ctx->stream));
    {
        PicProfScope ps(ctx, PIC_K_PUSH);
        size_t n = s.n_upper ? s.n_upper : 1;
        const unsigned nb = xcd_blocks1(n, BLOCK * PUSH_ILP);
        if (a.s.mode == PIC_SURF_ABSORB)
            k_push<true, PUSH_ILP, false><<<nb, BLOCK, 0, ctx->stream>>>(g, s.buf[s.cur], ctx->ef, a, s.d_n,
                                                                         s.count, s.rank);
        else
            k_push<true, PUSH_ILP, true><<<nb, BLOCK, 0, ctx->stream>>>(g, s.buf[s.cur], ctx->ef, a, s.d_n,
                                                                        s.count, s.rank);
    }
    PIC_HIP(ctx, hipGetLastError());
    s.hist_valid = true;
    s.sorted = false;
    s.two_part = false;
    s.tail_ok = false;  // cells changed
    return after_emission(ctx, s, a);
}

int pic_histogram(pic_ctx* ctx, int sid) {
    Species& s = ctx->sp[sid];
    const GridC& g = ctx->g;
    PIC_HIP(ctx, hipMemsetAsync(s.count, 0, ((size_t)g.ncells + 1) * 4, ctx->stream));
    {
        PicProfScope ps(ctx, PIC_K_SORT);
        size_t n = s.n_upper ? s.n_upper : 1;
        k_hist<<<xcd_blocks1(n, BLOCK), BLOCK, 0, ctx->stream>>>(g, s.buf[s.cur], s.d_n, s.count,
                                                              s.rank, 0u);
    }
    PIC_HIP(ctx, hipGetLastError());
    s.hist_valid = true;
    return PIC_OK;
}

// exclusive scan of the species' cell counts into out (the offsets; for a
// whole-species sort also the live count into d_n[0] and d_n[3])
static void scan_counts(pic_ctx* ctx, Species& s, uint32_t* out, bool whole) {
    const GridC& g = ctx->g;
    int M = g.ncells + 1;
    int nb = (M + SCAN_TILE - 1) / SCAN_TILE;
    PicProfScope ps(ctx, PIC_K_SORT);
    k_scan_reduce<<<nb, BLOCK, 0, ctx->stream>>>(s.count, M, ctx->scan_tmp);
    k_scan_top<<<1, BLOCK, 0, ctx->stream>>>(ctx->scan_tmp, nb);
    k_scan_down<<<nb, BLOCK, 0, ctx->stream>>>(s.count, M, ctx->scan_tmp, out, g.ncells,
                                               whole ? s.d_n : nullptr);
}

// pic_sort for scrambled input: the MSD passes above, ping-ponging between
// the two buffer sets; the sorted particles, offsets and live count as a
// counting sort leaves them (no ranks: the histogram is not valid after).
constexpr size_t PART_MIN = 1u << 20;  // particles; fewer fit in cache anyway
static int sort_partitioned(pic_ctx* ctx, Species& s) {
    const GridC& g = ctx->g;
    const size_t n = s.n_upper;
    const uint32_t kmax = (uint32_t)g.ncells;  // the removed particles' key
    int shift = 0;
    while ((kmax >> shift) + 1 > (uint32_t)P1_B) shift++;
    const int nbk = (int)(kmax >> shift) + 1;
    const size_t ntile = (n + PART_TILE - 1) / PART_TILE, M = (size_t)nbk * ntile;
    // segmented levels until a bucket spans at most FINE_CELLS keys and, at
    // the mean density, holds two thirds of FINE_CAP particles; each level
    // splits a bucket into up to PART_B (steps of at most 6 bits)
    int fine = FINE_SHIFT;
    while (fine > 0 && (double)n * (double)(1u << fine) > (2.0 / 3.0) * FINE_CAP * ((double)kmax + 1)) fine--;
    int steps[8], levels = 0;
    size_t nseg_max = nbk;
    for (int sh = shift; sh > fine;) {
        const int st = std::min(6, sh - fine);
        steps[levels++] = st;
        nseg_max <<= st;
        sh -= st;
    }
    const size_t nbound = nseg_max + 1;
    const size_t ntmax = ntile + nseg_max;  // tiles of any level
    const size_t nsum = M / SCAN_TILE + 2;
    const size_t words = M + (M + 1) + nsum + 2 * nbound + (nseg_max + 1) + ntmax * PART_B;
    if (M >= (1ull << 31) || ntmax * PART_B >= (1ull << 32))
        return pic_fail(ctx, PIC_ERR_CAPACITY, "too many particles to partition");
    if (ctx->part_words < words) {
        if (ctx->part) PIC_HIP(ctx, hipFree(ctx->part));
        ctx->part = nullptr;
        ctx->part_words = 0;
        PIC_HIP(ctx, hipMalloc(&ctx->part, words * 4));
        ctx->part_words = words;
    }
    uint32_t* pcount = ctx->part;
    uint32_t* poff = pcount + M;  // M + 1
    uint32_t* psum = poff + M + 1;
    uint32_t* bnd[2] = {psum + nsum, psum + nsum + nbound};
    uint32_t* tbase = bnd[1] + nbound;
    uint32_t* cnt2 = tbase + nseg_max + 1;
    int src = s.cur;
    {
        PicProfScope ps(ctx, PIC_K_SORT);
        k_part_count<<<(unsigned)ntile, PART_BLOCK, 0, ctx->stream>>>(g, s.buf[src].cell, s.d_n, shift, nbk,
                                                                     (uint32_t)ntile, pcount);
        const int nb = (int)((M + SCAN_TILE - 1) / SCAN_TILE);
        k_scan_reduce<<<nb, BLOCK, 0, ctx->stream>>>(pcount, (int)M, psum);
        k_scan_top<<<1, BLOCK, 0, ctx->stream>>>(psum, nb);
        k_scan_down<<<nb, BLOCK, 0, ctx->stream>>>(pcount, (int)M, psum, poff, -1, nullptr);
        k_part_bounds<<<(unsigned)((nbk + 1 + 255) / 256), 256, 0, ctx->stream>>>(poff, nbk, (uint32_t)ntile, bnd[0]);
    }
    {
        PicProfScope ps(ctx, PIC_K_SCATTER);
        k_part_move<<<(unsigned)ntile, PART_BLOCK, 0, ctx->stream>>>(g, s.buf[src], s.buf[src ^ 1], s.d_n, shift,
                                                                    nbk, (uint32_t)ntile, poff);
    }
    src ^= 1;
    int nseg = nbk, cur_b = 0;
    for (int l = 0; l < levels; l++) {
        const int st = steps[l];
        shift -= st;
        const unsigned tiles = (unsigned)(ntile + (size_t)nseg);
        {
            PicProfScope ps(ctx, PIC_K_SORT);
            k_seg_tiles<<<1, PART_BLOCK, 0, ctx->stream>>>(bnd[cur_b], nseg, tbase);
            k_seg_count<<<tiles, PART_BLOCK, 0, ctx->stream>>>(g, s.buf[src].cell, bnd[cur_b], tbase, nseg, shift,
                                                               st, cnt2);
            k_seg_scan<<<(unsigned)nseg, PART_B, 0, ctx->stream>>>(bnd[cur_b], tbase, nseg, st, cnt2,
                                                                   bnd[cur_b ^ 1]);
        }
        {
            PicProfScope ps(ctx, PIC_K_SCATTER);
            k_seg_move<<<tiles, PART_BLOCK, 0, ctx->stream>>>(g, s.buf[src], s.buf[src ^ 1], bnd[cur_b], tbase,
                                                              nseg, shift, st, cnt2);
        }
        src ^= 1;
        cur_b ^= 1;
        nseg <<= st;
    }
    PIC_HIP(ctx, hipMemsetAsync(s.count, 0, ((size_t)g.ncells + 1) * 4, ctx->stream));
    {
        PicProfScope ps(ctx, PIC_K_SCATTER);
        k_fine_sort<<<(unsigned)nseg, PART_BLOCK, 0, ctx->stream>>>(g, s.buf[src], s.buf[src ^ 1], bnd[cur_b], shift,
                                                                    s.off, s.count, s.rank, s.d_n);
    }
    PIC_HIP(ctx, hipGetLastError());
    s.cur = src ^ 1;
    s.hist_valid = false;
    return PIC_OK;
}

extern "C" int pic_sort(pic_ctx* ctx, int sid) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    Species& s = ctx->sp[sid];
    const GridC& g = ctx->g;
    const size_t n = s.n_upper ? s.n_upper : 1;
    if (s.scrambled && n >= PART_MIN) {
        rc = sort_partitioned(ctx, s);
        if (rc) return rc;
    } else {
        if (!s.hist_valid) {
            rc = pic_histogram(ctx, sid);
            if (rc) return rc;
        }
        scan_counts(ctx, s, s.off, true);
        {
            PicProfScope ps(ctx, PIC_K_SCATTER);
            k_scatter<<<xcd_blocks1((n + SC_V - 1) / SC_V, BLOCK), BLOCK, 0, ctx->stream>>>(
                g, s.buf[s.cur], s.buf[s.cur ^ 1], s.rank, s.off, s.d_n, (uint32_t)s.cap, 0u);
        }
        PIC_HIP(ctx, hipGetLastError());
        s.cur ^= 1;
    }
    s.scrambled = false;
    s.sorted = true;
    s.hist_valid = false;
    s.two_part = false;
    s.tail_ok = true;  // all sorted: an empty appended part
    return PIC_OK;
}

// Index the particles appended since the last whole-species sort (indices
// [bulk_n, n_upper), both exact: DSMC births, nothing removed) by cell: toff
// holds the appended part's per-cell offsets, tidx their positions in cell
// order.  The particles do not move; the species is then a sorted part with
// offsets off and an indexed appended part (a cell's pick pool: its sorted
// range, then its index entries).  An index that covers [bulk_n, t_end)
// takes only [t_end, n_upper) in; the order inside a cell is the index's
// (earlier appends first, then atomic order).
int pic_sort_tail(pic_ctx* ctx, int sid) {
    Species& s = ctx->sp[sid];
    const GridC& g = ctx->g;
    const size_t n = s.n_upper;
    if (n < s.bulk_n) return pic_fail(ctx, PIC_ERR_STATE, "appended part before the sorted part");
    const bool more = s.two_part && s.t_end >= s.bulk_n && s.t_end <= n;
    const size_t start = more ? s.t_end : s.bulk_n;
    if (n == start) {  // nothing appended since (e.g. only ionizing collisions)
        if (!more) s.two_part = false;  // the sorted part alone
        return PIC_OK;
    }
    const size_t words = ((size_t)g.ncells + 2) * 4, iw = (s.cap + 63) / 64 * 64 * 4;
    if (!s.toff) PIC_HIP(ctx, hipMalloc(&s.toff, words));
    if (!s.tidx) PIC_HIP(ctx, hipMalloc(&s.tidx, iw));
    PIC_HIP(ctx, hipMemsetAsync(s.count, 0, ((size_t)g.ncells + 1) * 4, ctx->stream));
    const size_t cnt = n - start;
    {
        PicProfScope ps(ctx, PIC_K_SORT);
        k_hist<<<xcd_blocks1(cnt, BLOCK), BLOCK, 0, ctx->stream>>>(g, s.buf[s.cur], s.d_n, s.count,
                                                                          s.rank, (uint32_t)start);
    }
    if (!more) {
        scan_counts(ctx, s, s.toff, false);
        PicProfScope ps(ctx, PIC_K_SCATTER);
        k_tail_index<<<grid_for(cnt, BLOCK), BLOCK, 0, ctx->stream>>>(g, s.buf[s.cur].cell, s.rank, s.toff,
                                                                     s.d_n, (uint32_t)start, s.tidx);
    } else {
        if (!s.boff) PIC_HIP(ctx, hipMalloc(&s.boff, words));
        if (!s.toff_n) PIC_HIP(ctx, hipMalloc(&s.toff_n, words));
        if (!s.tidx_n) PIC_HIP(ctx, hipMalloc(&s.tidx_n, iw));
        scan_counts(ctx, s, s.boff, false);
        PicProfScope ps(ctx, PIC_K_SCATTER);
        k_tail_merge<<<grid_for((size_t)g.ncells + 2, BLOCK), BLOCK, 0, ctx->stream>>>(
            s.toff, s.boff, g.ncells, s.tidx, s.toff_n, s.tidx_n);
        k_tail_index_more<<<grid_for(cnt, BLOCK), BLOCK, 0, ctx->stream>>>(
            g, s.buf[s.cur].cell, s.rank, s.toff, s.boff, s.d_n, (uint32_t)start, s.tidx_n);
        std::swap(s.toff, s.toff_n);
        std::swap(s.tidx, s.tidx_n);
    }
    PIC_HIP(ctx, hipGetLastError());
    s.two_part = true;
    s.t_end = n;
    s.hist_valid = false;
    return PIC_OK;
}

// A two-part species (pic_sort_tail just run) sorted whole by joining its
// parts: both stream once, where a counting sort scatters the appended part
// over the whole array.
int pic_join_tail(pic_ctx* ctx, int sid) {
    Species& s = ctx->sp[sid];
    const GridC& g = ctx->g;
    if (!s.two_part) return pic_fail(ctx, PIC_ERR_STATE, "no indexed appended part to join");
    const size_t nb = s.bulk_n, nt = s.n_upper - s.bulk_n;
    {
        PicProfScope ps(ctx, PIC_K_SCATTER);
        if (nb)
            k_join_sorted<<<xcd_blocks1(nb, BLOCK), BLOCK, 0, ctx->stream>>>(g, s.buf[s.cur], s.buf[s.cur ^ 1],
                                                                             s.toff, (uint32_t)nb);
        k_join_tail<<<xcd_blocks1(nt ? nt : 1, BLOCK), BLOCK, 0, ctx->stream>>>(
            g, s.buf[s.cur], s.buf[s.cur ^ 1], s.off, s.toff, s.tidx);
        k_join_off<<<grid_for((size_t)g.ncells + 1, BLOCK), BLOCK, 0, ctx->stream>>>(s.off, s.toff, g.ncells,
                                                                                    s.d_n);
    }
    PIC_HIP(ctx, hipGetLastError());
    s.cur ^= 1;
    s.sorted = true;
    s.scrambled = false;
    s.hist_valid = false;
    s.two_part = false;
    s.tail_ok = true;
    return PIC_OK;
}

// pic_push then pic_sort in two passes over the particles: the first pushes
// without writing them back and builds the sort's histogram (keeping each
// run's first slot in its cell), the second pushes again from the same
// inputs and writes every particle to its sorted position.  Same result as
// the two calls (the order inside a cell follows atomic order in both),
// about 25% less particle traffic, and no atomics in the second pass.
extern "C" int pic_push_sort(pic_ctx* ctx, int sid, double dt) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    Species& s = ctx->sp[sid];
    const GridC& g = ctx->g;
    // random-order input (a load or an upload): sorted first, so that both
    // passes see a wave's particles in a few cells (same result: the push is
    // per particle and the output is sorted either way)
    if (s.scrambled && s.n_upper >= PART_MIN) {
        rc = pic_sort(ctx, sid);
        if (rc) return rc;
    }
    PushArgs a = push_args(ctx, s, dt);  // emission in the second pass only
    const size_t n = s.n_upper ? s.n_upper : 1;
    PIC_HIP(ctx, hipMemsetAsync(s.count, 0, ((size_t)g.ncells + 1) * 4, ctx->stream));
    {
        PicProfScope ps(ctx, PIC_K_PUSH);
        const unsigned nb = xcd_blocks1(n, BLOCK * PUSH_SORT_ILP);
        if (a.s.mode == PIC_SURF_ABSORB)
            k_push<false, PUSH_SORT_ILP, false><<<nb, BLOCK, 0, ctx->stream>>>(g, s.buf[s.cur], ctx->ef, a,
                                                                               s.d_n, s.count, s.rank);
        else
            k_push<false, PUSH_SORT_ILP, true><<<nb, BLOCK, 0, ctx->stream>>>(g, s.buf[s.cur], ctx->ef, a,
                                                                              s.d_n, s.count, s.rank);
    }
    scan_counts(ctx, s, s.off, true);
    a.s.emit = 1;
    {
        PicProfScope ps(ctx, PIC_K_SCATTER);
        const unsigned nb2 = xcd_blocks1(n, BLOCK * PUSH_SCATTER_ILP);
        if (a.s.mode == PIC_SURF_ABSORB)
            k_push_scatter<false, PUSH_SCATTER_ILP><<<nb2, BLOCK, 0, ctx->stream>>>(
                g, s.buf[s.cur], s.buf[s.cur ^ 1], ctx->ef, a, s.rank, s.off, s.d_n, (uint32_t)s.cap);
        else
            k_push_scatter<true, 1><<<xcd_blocks1(n, BLOCK), BLOCK, 0, ctx->stream>>>(
                g, s.buf[s.cur], s.buf[s.cur ^ 1], ctx->ef, a, s.rank, s.off, s.d_n, (uint32_t)s.cap);
    }
    PIC_HIP(ctx, hipGetLastError());
    s.cur ^= 1;
    s.sorted = true;
    s.scrambled = false;
    s.hist_valid = false;
    s.two_part = false;
    s.tail_ok = true;
    return after_emission(ctx, s, a);
}

extern "C" int pic_deposit(pic_ctx* ctx, int sid, int flags) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    Species& s = ctx->sp[sid];
    const GridC& g = ctx->g;
    if (!(flags & (PIC_DEPOSIT_DENSITY | PIC_DEPOSIT_MOMENTS))) return PIC_ERR_ARG;
    if (!s.sorted) {
        rc = pic_sort(ctx, sid);
        if (rc) return rc;
    }
    size_t nn = (size_t)g.nn;
    int want_den = (flags & PIC_DEPOSIT_DENSITY) != 0;
    bool mom = (flags & PIC_DEPOSIT_MOMENTS) != 0;
    // corner sums: 8 per cell for the density, 7 x 8 for the moments
    const size_t corner_floats = (mom ? 7 * 8 : 8) * (size_t)g.ncells;
    if (ctx->dep_corner_floats < corner_floats) {
        PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        (void)hipFree(ctx->dep_corner);
        ctx->dep_corner = nullptr;
        ctx->dep_corner_floats = 0;
        PIC_HIP(ctx, hipMalloc(&ctx->dep_corner, corner_floats * sizeof(float)));
        ctx->dep_corner_floats = corner_floats;
    }
    if (mom && !s.mom) {
        PIC_HIP(ctx, hipMalloc(&s.mom, 7 * nn * 8));
        PIC_HIP(ctx, hipMemsetAsync(s.mom, 0, 7 * nn * 8, ctx->stream));
    }
    {
        PicProfScope ps(ctx, PIC_K_DEPOSIT);
        if (mom) {
            k_mom_cells<<<grid_for((size_t)g.ncells, BLOCK), BLOCK, 0, ctx->stream>>>(
                g, s.buf[s.cur], s.off, ctx->dep_corner);
            k_mom_nodes<<<grid_for(nn, BLOCK), BLOCK, 0, ctx->stream>>>(g, ctx->dep_corner, s.mom,
                                                                        want_den ? s.den : nullptr);
        } else {
            pic_density_launch(ctx, sid, ctx->dep_corner);
        }
    }
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

// the density of a sorted species through corner sums in `corner` (8
// floats per cell, every one written before it is read): the two launches,
// queued on the context's stream, no allocation, no wait
void pic_density_launch(pic_ctx* ctx, int sid, float* corner) {
    Species& s = ctx->sp[sid];
    const GridC& g = ctx->g;
    k_dep_cells<<<grid_for((size_t)g.ncells, BLOCK), BLOCK, 0, ctx->stream>>>(g, s.buf[s.cur], s.off, corner);
    k_dep_nodes<<<grid_for((size_t)g.nn, BLOCK), BLOCK, 0, ctx->stream>>>(g, corner, s.den);
}

extern "C" int pic_clear_samples(pic_ctx* ctx, int sid) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    Species& s = ctx->sp[sid];
    if (s.mom) PIC_HIP(ctx, hipMemsetAsync(s.mom, 0, 7 * (size_t)ctx->g.nn * 8, ctx->stream));
    return PIC_OK;
}

extern "C" int pic_charge_density(pic_ctx* ctx) {
    if (!ctx) return PIC_ERR_ARG;
    RhoArgs a;
    a.ns = 0;
    for (auto& s : ctx->sp) {
        if (s.charge == 0) continue;  // World.cpp:197
        a.den[a.ns] = s.den;
        a.q[a.ns] = s.charge;
        a.ns++;
    }
    {
        PicProfScope ps(ctx, PIC_K_RHO);
        k_rho<<<grid_for((size_t)ctx->g.nn, BLOCK), BLOCK, 0, ctx->stream>>>(ctx->g.nn, a,
                                                                            ctx->rho);
    }
    PIC_HIP(ctx, hipGetLastError());
    return PIC_OK;
}

extern "C" int pic_load_box_thermal(pic_ctx* ctx, int sid, const double centre[3],
                                    const double sides[3], double num_den, double T,
                                    uint32_t seed, int mode, size_t* n_loaded) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    if (!centre || !sides || (mode != 0 && mode != 1)) return PIC_ERR_ARG;
    Species& s = ctx->sp[sid];
    double box_vol = sides[0] * sides[1] * sides[2];
    double num_micro = num_den * box_vol;
    size_t num_macro = (size_t)(num_micro / s.mpw0);  // Species.cpp:565-567
    if (num_macro < 1)
        return pic_fail(ctx, PIC_ERR_ARG, "number of macroparticles less than 1");
    uint32_t hn[4];
    PIC_HIP(ctx, hipMemcpyAsync(hn, s.d_n, sizeof(hn), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    size_t base = hn[0];
    if (base + num_macro > s.cap)
        return pic_fail(ctx, PIC_ERR_CAPACITY, "loader exceeds species capacity");
    LoadArgs a;
    for (int d = 0; d < 3; d++) {
        a.xmin[d] = centre[d] - sides[d] * 0.5;
        a.xmax[d] = centre[d] + sides[d] * 0.5;
    }
    a.vth = mode == 0 ? sqrt(2 * KB * T / s.mass) : sqrt(KB * T / s.mass);
    a.rewind = s.charge / s.mass * (0.5 * ctx->dt);
    a.seed = seed;
    a.seq = s.seq;
    a.mode = mode;
    a.w = (float)s.mpw0;
    {
        PicProfScope ps(ctx, PIC_K_LOAD);
        k_load<<<grid_for(num_macro, BLOCK), BLOCK, 0, ctx->stream>>>(ctx->g, s.buf[s.cur],
                                                                      ctx->ef, a, base, num_macro);
    }
    PIC_HIP(ctx, hipGetLastError());
    hn[0] = (uint32_t)(base + num_macro);
    PIC_HIP(ctx, hipMemcpyAsync(s.d_n, hn, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s.n_upper = base + num_macro;
    s.hist_valid = false;
    s.sorted = false;
    s.two_part = false;
    s.tail_ok = false;
    s.scrambled = true;  // uniform positions in thread order
    if (n_loaded) *n_loaded = num_macro;
    return PIC_OK;
}

extern "C" int pic_merge(pic_ctx* ctx, int sid, uint32_t seed, pic_merge_stats* stats) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    Species& s = ctx->sp[sid];
    const GridC& g = ctx->g;
    if (!s.sorted || s.two_part) {  // one contiguous range per cell
        rc = pic_sort(ctx, sid);
        if (rc) return rc;
    }
    const size_t n_before = s.n_upper;
    rc = pic_ensure_pin(ctx, 64);
    if (rc) return rc;
    unsigned long long* d = reinterpret_cast<unsigned long long*>(ctx->scratch);  // 3 counters
    PIC_HIP(ctx, hipMemsetAsync(d, 0, 3 * sizeof(unsigned long long), ctx->stream));
    MergeArgs a;
    a.seed = seed;
    a.call = s.merge_calls++;
    a.rewind = s.charge != 0.0 ? s.charge / s.mass * (0.5 * ctx->dt) : 0.0;
    k_merge<<<(unsigned)g.ncells, MERGE_BLOCK, 0, ctx->stream>>>(g, s.buf[s.cur], s.off, ctx->ef, a, d);
    PIC_HIP(ctx, hipGetLastError());
    s.sorted = false;  // removed members leave at the compaction below
    s.hist_valid = false;
    s.tail_ok = false;
    s.two_part = false;
    rc = pic_sort(ctx, sid);
    if (rc) return rc;
    unsigned long long* h = reinterpret_cast<unsigned long long*>(ctx->pin);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(ctx->pin + 4);
    PIC_HIP(ctx, hipMemcpyAsync(h, d, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipMemcpyAsync(cnt, s.d_n, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s.n_upper = cnt[0];
    if (stats) {
        stats->bins = (long long)h[0];
        stats->merged = (long long)h[1];
        stats->removed = (long long)n_before - (long long)cnt[0];
    }
    if (h[2])
        return pic_fail(ctx, PIC_ERR_CAPACITY, "merge: a cell holds more than 8192 particles");
    return PIC_OK;
}
