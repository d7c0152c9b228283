// pic_collide.hip -- equal-weight DSMC (DSMC_MEX, Interactions.cpp:143-284)
// and beam sources (ColdBeamSource / WarmBeamSource, Source.cpp:3-185) on
// gfx950.
//
// DSMC: one thread owns one cell and runs its No-Time-Counter groups in
// order over the cell-sorted SoA arrays of one or two species; velocities
// are updated in place in FP64 arithmetic; the step maximum of sigma v_rel
// is reduced over striped per-workgroup slots.  Draws per (cell, group):
// Philox slots 0 accept, 1 cos(chi), 2 azimuth, 3 first pick, 4.. the second
// pick and its redraws.
//
// Sources: one thread per injected particle, appended behind the species'
// live particles; addParticle's rejected points are stored DEAD and leave
// at the next sort.
#include <algorithm>
#include <cmath>

#include "pic_internal.h"

using namespace pic;

namespace {

constexpr int BLOCK = 128;

struct DsmcArgs {
    double m1, m2, sum_m, m_red;
    double c0, c1, c2, c3;  // Bird's VHS parameters (Interactions.cpp:153-156)
    double mpw0, dt, dv;    // dv = dx dy dz (World::getCellVolume, World.cpp:101-103)
    uint32_t seed, step;
    int one;                // one species (distinct picks)
};

// Interactions.cpp:175-177
__device__ inline double vhs_sigma(const DsmcArgs& a, double v_rel) {
    return PI_REF * a.c0 * a.c0 * pow(a.c2 / (v_rel * v_rel), a.c1 - 0.5) / a.c3;
}

__global__ void __launch_bounds__(BLOCK)
    k_dsmc(GridC g, DsmcArgs a, SpeciesBuf b1, const uint32_t* __restrict__ off1, SpeciesBuf b2,
           const uint32_t* __restrict__ off2, const double* __restrict__ sv,
           unsigned long long* __restrict__ part) {
    unsigned long long* stats = part + (blockIdx.x % NSTRIPE) * SSTRIDE;
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    double tmp_max = 0;
    unsigned long long ncoll = 0, ngr = 0;
    if (c < g.ncells) {
        const uint32_t p1 = off1[c], p2 = off2[c];
        const int N1 = (int)(off1[c + 1] - p1), N2 = (int)(off2[c + 1] - p2);
        const bool act = a.one ? N1 >= 2 : (N1 >= 1 && N2 >= 1);
        const double svmax = sv[0];
        // Interactions.cpp:195, :241 in the same operation order (an exact
        // tie at .5 then rounds the same way)
        const int ng = act ? (int)(0.5 * (double)N1 * (double)N2 * a.mpw0 * svmax * a.dt / a.dv + 0.5) : 0;
        for (int gi = 0; gi < ng; gi++) {
            Draws dr;
            dr.init((uint32_t)c, (uint32_t)gi, a.seed, a.step, TAG_DSMC);
            ngr++;
            int slot = 3;
            const int ia = (int)(dr(slot++) * (double)N1);
            int ib = (int)(dr(slot++) * (double)N2);
            while (a.one && ib == ia) ib = (int)(dr(slot++) * (double)N2);
            const uint32_t qa = p1 + (uint32_t)ia, qb = p2 + (uint32_t)ib;
            double va[3] = {b1.v[0][qa], b1.v[1][qa], b1.v[2][qa]};
            double vb[3] = {b2.v[0][qb], b2.v[1][qb], b2.v[2][qb]};
            double vr[3] = {va[0] - vb[0], va[1] - vb[1], va[2] - vb[2]};
            const double v_rel = sqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
            const double s = vhs_sigma(a, v_rel) * v_rel;
            if (s > tmp_max) tmp_max = s;
            if (!(s / svmax > dr(0))) continue;
            ncoll++;
            // Interactions.cpp:264-284: isotropic scattering in the centre of mass
            const double inv = 1.0 / a.sum_m;
            double cm[3];
#pragma unroll
            for (int d = 0; d < 3; d++) cm[d] = (va[d] * a.m1 + vb[d] * a.m2) * inv;
            const double cos_ksi = 2 * dr(1) - 1;
            const double sin_ksi = sqrt(1 - cos_ksi * cos_ksi);
            const double eps = 2 * PI_REF * dr(2);
            const double r[3] = {v_rel * cos_ksi, v_rel * sin_ksi * cos(eps), v_rel * sin_ksi * sin(eps)};
            const double f2 = a.m2 / a.sum_m, f1 = a.m1 / a.sum_m;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                b1.v[d][qa] = (float)(cm[d] + r[d] * f2);
                b2.v[d][qb] = (float)(cm[d] - r[d] * f1);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        tmp_max = fmax(tmp_max, __shfl_xor(tmp_max, o, 64));
        ncoll += __shfl_xor(ncoll, o, 64);
        ngr += __shfl_xor(ngr, o, 64);
    }
    if ((threadIdx.x & 63) == 0 && ngr) {
        atomicAdd(&stats[0], ncoll);
        atomicAdd(&stats[1], ngr);
        atomicMax(&stats[2], (unsigned long long)__double_as_longlong(tmp_max));
    }
}

// stripes summed; sigma v_rel max <- the step's maximum when anything
// collided (Interactions.cpp:224-227)
__global__ void k_dsmc_finish(double* sv, const unsigned long long* part, unsigned long long* stats) {
    unsigned long long t0 = 0, t1 = 0, mb = 0;
    for (int s = 0; s < NSTRIPE; s++) {
        t0 += part[s * SSTRIDE];
        t1 += part[s * SSTRIDE + 1];
        mb = part[s * SSTRIDE + 2] > mb ? part[s * SSTRIDE + 2] : mb;
    }
    stats[0] = t0;
    stats[1] = t1;
    sv[1] = __longlong_as_double((long long)mb);
    if (t0 > 0) sv[0] = sv[1];
}

struct SrcArgs {
    int ax, plus, t0, t1;
    double x_face;      // the face's coordinate (the + face one cell inside)
    double lo[3], L[3]; // the face's origin and extents
    double v_drift;
    double vth;         // sqrt(2 k T / m) of sampleVth; 0: cold beam
    double rewind;      // (q/m) * 0.5 dt
    uint32_t seed, call;
    float w;
};

__device__ inline bool in_rect_d(const GridC& g, const double* lc) {
    for (int r = 0; r < g.nrect; r++) {
        bool in = true;
#pragma unroll
        for (int d = 0; d < 3; d++)
            if (lc[d] < g.rlo[r][d] || lc[d] > g.rhi[r][d]) in = false;
        if (in) return true;
    }
    return false;
}

// Source.cpp:38-97 (cold), :111-180 (warm), then addParticle
// (Species.cpp:420-434): the particle's draws are slots 0-10 (warm
// sampleV3th, Species.cpp:855-869) and 11-12 (the two face coordinates)
__global__ void __launch_bounds__(256)
    k_source(GridC g, SpeciesBuf b, const float4* __restrict__ E, SrcArgs a, size_t base, size_t count) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    Draws dr;
    dr.init((uint32_t)i, 0u, a.seed, a.call, TAG_SRC);
    double vel[3] = {0, 0, 0};
    if (a.vth > 0) {
        double v[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double u1 = dr(3 * c), u2 = dr(3 * c + 1), u3 = dr(3 * c + 2);
            v[c] = a.vth * (u1 + u2 + u3 - 1.5);
        }
        const double sp = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        const double theta = 2 * PI_REF * dr(9);
        const double r = -1.0 + 2 * dr(10);
        const double s = sqrt(1 - r * r);
        vel[0] = sp * r;
        vel[1] = sp * (cos(theta) * s);
        vel[2] = sp * (sin(theta) * s);
    }
    vel[a.ax] += a.plus ? -a.v_drift : a.v_drift;
    double pos[3];
    pos[a.ax] = a.x_face;
    pos[a.t0] = a.lo[a.t0] + dr(11) * a.L[a.t0];
    pos[a.t1] = a.lo[a.t1] + dr(12) * a.L[a.t1];
    double lc[3];
    int ijk[3];
    const int lim[3] = {g.ci, g.cj, g.ck};
    float f[3] = {0, 0, 0};
    uint32_t cw = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        lc[d] = (pos[d] - g.x0[d]) * g.inv_dx[d];
        if (!(lc[d] >= 0.0)) cw = DEAD;
        ijk[d] = (int)lc[d];
        if (ijk[d] >= lim[d]) cw = DEAD;
        const float fr = (float)(lc[d] - ijk[d]);
        f[d] = fr < 1.0f ? fr : 0.99999994f;
    }
    if (cw != DEAD && g.nrect && in_rect_d(g, lc)) cw = DEAD;
    if (cw != DEAD) {
        cw = pack_cell(ijk[0], ijk[1], ijk[2]);
        if (a.rewind != 0.0) {
            // CIC gather of E at the particle (Field.h:201-232)
            const unsigned sj = (unsigned)g.ni, sk = (unsigned)g.ni * (unsigned)g.nj;
            const unsigned u = (unsigned)ijk[2] * sk + (unsigned)ijk[1] * sj + (unsigned)ijk[0];
            double e[3] = {0, 0, 0};
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int di = q & 1, dj = (q >> 1) & 1, dk = q >> 2;
                const double w = (di ? f[0] : 1.0 - f[0]) * (dj ? f[1] : 1.0 - f[1]) * (dk ? f[2] : 1.0 - f[2]);
                const float4 ev = E[u + di + dj * sj + dk * sk];
                e[0] += w * ev.x;
                e[1] += w * ev.y;
                e[2] += w * ev.z;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) vel[d] -= e[d] * a.rewind;
        }
    }
    const size_t o = base + i;
    b.cell[o] = cw;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        b.f[d][o] = f[d];
        b.v[d][o] = (float)vel[d];
    }
    b.w[o] = a.w;
}

__global__ void k_set_face(GridC g, double* __restrict__ phi, int face, double value) {
    const int ax = face / 2, plus = face & 1;
    const int n[3] = {g.ni, g.nj, g.nk};
    const int t0 = ax == 0 ? 1 : 0, t1 = ax == 2 ? 1 : 2;
    const long long m = (long long)blockIdx.x * 256 + threadIdx.x;
    if (m >= (long long)n[t0] * n[t1]) return;
    int ijk[3];
    ijk[ax] = plus ? n[ax] - 1 : 0;
    ijk[t0] = (int)(m % n[t0]);
    ijk[t1] = (int)(m / n[t0]);
    phi[((long long)ijk[2] * g.nj + ijk[1]) * g.ni + ijk[0]] = value;
}

}  // namespace

extern "C" int pic_dsmc_create(pic_ctx* ctx, int sid1, int sid2, int* id) {
    if (!ctx || !id) return PIC_ERR_ARG;
    int rc = pic_check_sid(ctx, sid1);
    if (!rc) rc = pic_check_sid(ctx, sid2);
    if (rc) return rc;
    if (sid1 != sid2 && ctx->sp[sid1].mpw0 != ctx->sp[sid2].mpw0)  // Interactions.cpp:160-162
        return pic_fail(ctx, PIC_ERR_ARG,
                        "species must have the same macroparticle weight for this algorithm to work properly.");
    Dsmc d;
    d.s1 = sid1;
    d.s2 = sid2;
    PIC_HIP(ctx, hipMalloc(&d.d_sv, 2 * sizeof(double)));
    PIC_HIP(ctx, hipMalloc(&d.d_part, (size_t)NSTRIPE * SSTRIDE * sizeof(unsigned long long)));
    PIC_HIP(ctx, hipMalloc(&d.d_stats, 2 * sizeof(unsigned long long)));
    const double h[2] = {1e-14, 0.0};  // Interactions.h:60
    PIC_HIP(ctx, hipMemcpyAsync(d.d_sv, h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->dsmc.push_back(d);
    *id = (int)ctx->dsmc.size() - 1;
    return PIC_OK;
}

extern "C" int pic_dsmc_apply(pic_ctx* ctx, int id, double dt, uint32_t seed, uint32_t step,
                              pic_dsmc_stats* stats) {
    if (!ctx || id < 0 || id >= (int)ctx->dsmc.size()) return PIC_ERR_ARG;
    Dsmc& D = ctx->dsmc[id];
    Species &A = ctx->sp[D.s1], &B = ctx->sp[D.s2];
    for (Species* s : {&A, &B})
        if (!s->sorted || s->two_part) {
            int rc = pic_sort(ctx, (int)(s - ctx->sp.data()));
            if (rc) return rc;
        }
    const GridC& g = ctx->g;
    DsmcArgs a;
    a.m1 = A.mass;
    a.m2 = B.mass;
    a.sum_m = A.mass + B.mass;
    a.m_red = A.mass * B.mass / (A.mass + B.mass);
    a.c0 = 4.07e-10;
    a.c1 = 0.77;
    a.c2 = 2 * KB * 273.15 / a.m_red;
    a.c3 = std::tgamma(2.5 - a.c1);
    a.mpw0 = A.mpw0;
    a.dt = dt;
    a.dv = g.dx[0] * g.dx[1] * g.dx[2];
    a.seed = seed;
    a.step = step;
    a.one = D.s1 == D.s2;
    PIC_HIP(ctx, hipMemsetAsync(D.d_part, 0, (size_t)NSTRIPE * SSTRIDE * sizeof(unsigned long long),
                                ctx->stream));
    {
        PicProfScope ps(ctx, PIC_K_MCC);
        k_dsmc<<<grid_for((size_t)g.ncells, BLOCK), BLOCK, 0, ctx->stream>>>(g, a, A.buf[A.cur], A.off,
                                                                             B.buf[B.cur], B.off, D.d_sv,
                                                                             D.d_part);
        k_dsmc_finish<<<1, 1, 0, ctx->stream>>>(D.d_sv, D.d_part, D.d_stats);
    }
    PIC_HIP(ctx, hipGetLastError());
    int rc = pic_ensure_pin(ctx, 64);
    if (rc) return rc;
    unsigned long long* hs = reinterpret_cast<unsigned long long*>(ctx->pin);
    double* hv = ctx->pin + 2;
    PIC_HIP(ctx, hipMemcpyAsync(hs, D.d_stats, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                ctx->stream));
    PIC_HIP(ctx, hipMemcpyAsync(hv, D.d_sv, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, pic_spin_wait(ctx->stream));
    // velocities changed in place: cells (and the sort) are unchanged
    if (stats) {
        stats->collisions = (long long)hs[0];
        stats->groups = (long long)hs[1];
        stats->sigma_v_rel_max = hv[0];
    }
    return PIC_OK;
}

extern "C" int pic_source_create(pic_ctx* ctx, int sid, int face, double v_drift, double den, double T,
                                 int* id) {
    int rc = pic_check_sid(ctx, sid);
    if (rc) return rc;
    if (!id || face < 0 || face > 5) return pic_fail(ctx, PIC_ERR_ARG, "Passed wrong face name.");
    const GridC& g = ctx->g;
    const int ax = face / 2, t0 = ax == 0 ? 1 : 0, t1 = ax == 2 ? 1 : 2;
    const int n[3] = {g.ni, g.nj, g.nk};
    Src s;
    s.sid = sid;
    s.face = face;
    s.v_drift = v_drift;
    s.den = den;
    s.T = T;
    s.dt = ctx->dt;
    s.num_micro = den * v_drift * (g.dx[t0] * (n[t0] - 1)) * (g.dx[t1] * (n[t1] - 1)) * ctx->dt;  // :19-26
    // World::addInlet (World.cpp:206-262): the face is Dirichlet at phi = 0
    for (int k = 0; k < g.nk; k++)
        for (int j = 0; j < g.nj; j++)
            for (int i = 0; i < g.ni; i++) {
                const int ijk[3] = {i, j, k};
                if (ijk[ax] == ((face & 1) ? n[ax] - 1 : 0))
                    ctx->node_type_in[((size_t)k * g.nj + j) * g.ni + i] = 2;
            }
    ctx->ntype_dirty = true;
    k_set_face<<<grid_for((size_t)n[t0] * n[t1], 256), 256, 0, ctx->stream>>>(g, ctx->phi, face, 0.0);
    PIC_HIP(ctx, hipGetLastError());
    ctx->src.push_back(s);
    *id = (int)ctx->src.size() - 1;
    return PIC_OK;
}

extern "C" int pic_source_sample(pic_ctx* ctx, int id, uint32_t seed, size_t* n_added) {
    if (!ctx || id < 0 || id >= (int)ctx->src.size()) return PIC_ERR_ARG;
    Src& S = ctx->src[id];
    Species& sp = ctx->sp[S.sid];
    const GridC& g = ctx->g;
    const uint32_t call = S.calls++;
    Draws dr;  // the count's draw, on the host (the same Philox stream)
    dr.init(0xFFFFFFFFu, 0xFFFFFFFFu, seed, call, TAG_SRC);
    const long long num_macro = (long long)(S.num_micro / sp.mpw0 + dr(0));
    if (n_added) *n_added = 0;
    if (num_macro <= 0) return PIC_OK;
    uint32_t hn[4];
    PIC_HIP(ctx, hipMemcpyAsync(hn, sp.d_n, sizeof(hn), hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t base = hn[0];
    if (base + (size_t)num_macro > sp.cap)
        return pic_fail(ctx, PIC_ERR_CAPACITY, "source exceeds species capacity");
    SrcArgs a;
    a.ax = S.face / 2;
    a.plus = S.face & 1;
    a.t0 = a.ax == 0 ? 1 : 0;
    a.t1 = a.ax == 2 ? 1 : 2;
    const int n[3] = {g.ni, g.nj, g.nk};
    for (int d = 0; d < 3; d++) {
        a.lo[d] = g.x0[d];
        a.L[d] = g.dx[d] * (n[d] - 1);
    }
    a.x_face = g.x0[a.ax] + (a.plus ? a.L[a.ax] - g.dx[a.ax] : 0.0);  // Source.cpp:47-48
    a.v_drift = S.v_drift;
    a.vth = S.T > 0 ? std::sqrt(2 * KB * S.T / sp.mass) : 0.0;
    a.rewind = sp.charge / sp.mass * (0.5 * ctx->dt);
    a.seed = seed;
    a.call = call;
    a.w = (float)sp.mpw0;
    {
        PicProfScope ps(ctx, PIC_K_LOAD);
        k_source<<<grid_for((size_t)num_macro, 256), 256, 0, ctx->stream>>>(g, sp.buf[sp.cur], ctx->ef, a, base,
                                                                            (size_t)num_macro);
    }
    PIC_HIP(ctx, hipGetLastError());
    hn[0] = (uint32_t)(base + (size_t)num_macro);
    PIC_HIP(ctx, hipMemcpyAsync(sp.d_n, hn, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    sp.n_upper = base + (size_t)num_macro;
    sp.hist_valid = false;
    sp.sorted = false;
    sp.two_part = false;
    sp.tail_ok = false;  // rejected points are DEAD: a whole sort removes them
    if (n_added) *n_added = (size_t)num_macro;
    return PIC_OK;
}
