// pic_mcc.hip -- variable-weight electron-neutral Monte-Carlo collisions with
// ionization (MC_MEX_Ionization, Interactions.cpp:476-980) on gfx950.
//
// One thread owns one cell and runs that cell's No-Time-Counter group loop
// sequentially, exactly as apply_vector_indexes does (Interactions.cpp:
// 627-748); cells are independent, so the grid of cells is the parallel
// dimension.  Particles are read from the cell-sorted SoA arrays of the three
// species; weights and velocities are updated in place; births (ions,
// electrons, split neutrals) are appended at the tail of each species through
// an atomic counter.  All collision arithmetic is FP64.
//
// Draws: Philox4x32-10, key (seed, step), counter (cell, group, slot/2, tag);
// slots 0 neutral pick, 1 electron pick, 2 accept, 3 ionize test, 4.. the
// scattering angles and the ejected energy (same order as the reference's
// sequential rnd() calls, Interactions.cpp:656-678, :895-944).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pic_internal.h"

using namespace pic;

namespace {

constexpr int BLOCK = 128;

struct MccArgs {
    double m_n, m_e, sum_m, m_red;
    double E_pre_rel;  // 0.5 m_red / q          (Interactions.cpp:501)
    double E_pre_ele;  // m_e 0.5 / q            (Interactions.cpp:123)
    double two_q_me;   // 2 q / m_e              (Interactions.cpp:124)
    double E_ion_eV, c0, c1, c2, B_inc;
    double dt, inv_dv;  // inv_dv = 1 / (dx dy dz) (Interactions.cpp:493-494)
    uint32_t seed, step;
    int nsig;
};

struct Sp {
    SpeciesBuf b;
    uint32_t* d_n;
    uint32_t cap;
};

// Interactions.cpp:541-558: std::map::lower_bound + linear interpolation
__device__ inline double sigma_el(const double* __restrict__ tE, const double* __restrict__ tV,
                                  int n, double E) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (tE[mid] < E)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo == 0) return tV[0];
    if (lo == n) return tV[n - 1];
    double x1 = tE[lo - 1], x2 = tE[lo], y1 = tV[lo - 1], y2 = tV[lo];
    return y1 + (E - x1) * (y2 - y1) / (x2 - x1);
}
// Interactions.cpp:559-566
__device__ inline double sigma_iz(const MccArgs& a, double E) {
    if (E <= a.E_ion_eV) return 0;
    return a.c0 * log(E / a.c1) / E * exp(-a.c2 / E);
}

__device__ inline void cross3(const double* x, const double* y, double* o) {
    o[0] = x[1] * y[2] - x[2] * y[1];
    o[1] = x[2] * y[0] - x[0] * y[2];
    o[2] = x[0] * y[1] - x[1] * y[0];
}

// Interactions.cpp:845-870 (the basis i x u, u x (i x u) is not normalised
// in the reference; kept as is)
__device__ inline void new_vel_electron(const MccArgs& a, double* vel, double E, const double* u,
                                        double U1, double U2) {
    double cos_ksi = (2 + E - 2 * pow(1 + E, U1)) / E;
    double sin_ksi = sqrt(1 - cos_ksi * cos_ksi);
    double phi = 2 * PI_REF * U2;
    double v_mag = sqrt(E * a.two_q_me);
    const double ix[3] = {1, 0, 0};
    double ic[3], uic[3];
    cross3(ix, u, ic);
    cross3(u, ic, uic);
    double sp = sin(phi), cp = cos(phi);
#pragma unroll
    for (int d = 0; d < 3; d++) vel[d] = (u[d] * cos_ksi + ic[d] * sin_ksi * sp + uic[d] * sin_ksi * cp) * v_mag;
}

// Interactions.cpp:885-947 (IONIZE_1).  vn is not modified.
__device__ inline bool collide(const MccArgs& a, Draws& dr, const double* vn, double* ve,
                               double* vnew, double sigma_coll) {
    double vr[3] = {vn[0] - ve[0], vn[1] - ve[1], vn[2] - ve[2]};
    double vrm = sqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
    double E_rel_J = 0.5 * a.m_red * vrm * vrm;
    double P = sigma_iz(a, E_rel_J / QE) / sigma_coll;
    if (dr(3) <= P) {
        double vem = sqrt(ve[0] * ve[0] + ve[1] * ve[1] + ve[2] * ve[2]);
        double E_inc = vem * vem * a.E_pre_ele;
        if (E_inc < a.E_ion_eV) return false;  // Interactions.cpp:901-905
        double E_ej = 10.0 * tan(dr(4) * atan((E_inc - a.E_ion_eV) / (2 * a.B_inc)));
        double E_sc = E_inc - a.E_ion_eV - E_ej;
        if (E_sc < 0) E_sc = 0.000001;
        double inv = 1.0 / vem;
        double u[3] = {ve[0] * inv, ve[1] * inv, ve[2] * inv};
        double U5 = dr(5), U6 = dr(6);
        new_vel_electron(a, ve, E_sc, u, U5, U6);
        double U7 = dr(7), U8 = dr(8);
        new_vel_electron(a, vnew, E_ej, u, U7, U8);
        return true;
    }
    double inv = 1.0 / a.sum_m;
    double cm[3];
#pragma unroll
    for (int d = 0; d < 3; d++) cm[d] = (vn[d] * a.m_n + ve[d] * a.m_e) * inv;
    double cos_ksi = 2 * dr(4) - 1;
    double sin_ksi = sqrt(1 - cos_ksi * cos_ksi);
    double eps = 2 * PI_REF * dr(5);
    double v2[3] = {vrm * cos_ksi, vrm * sin_ksi * cos(eps), vrm * sin_ksi * sin(eps)};
    double f = a.m_n / a.sum_m;
#pragma unroll
    for (int d = 0; d < 3; d++) ve[d] = cm[d] - v2[d] * f;
    return false;
}

// Births are staged in LDS per workgroup (slots from LDS atomics, which
// return within the CU) and appended to the species' tails once per
// workgroup at the end: a returning global atomic per birth left the wave
// waiting on a device-wide round trip up to three times per NTC group.
// Staging slots per species (neutral splits dominate); a full buffer falls
// back to a global slot for that birth.
struct Birth {
    uint32_t cell;
    float f[3], v[3], w;
};
constexpr int BIRTH_CAP_N = 512, BIRTH_CAP_I = 64, BIRTH_CAP_E = 96;
struct BirthStage {
    Birth b[BIRTH_CAP_N + BIRTH_CAP_I + BIRTH_CAP_E];
    uint32_t count[3], base[3];
};
__device__ inline int birth_cap(int k) { return k == 0 ? BIRTH_CAP_N : (k == 1 ? BIRTH_CAP_I : BIRTH_CAP_E); }
__device__ inline int birth_off(int k) { return k == 0 ? 0 : (k == 1 ? BIRTH_CAP_N : BIRTH_CAP_N + BIRTH_CAP_I); }

__device__ inline void store_birth(Sp& s, uint32_t idx, const Birth& b) {
    s.b.cell[idx] = b.cell;
    s.b.f[0][idx] = b.f[0];
    s.b.f[1][idx] = b.f[1];
    s.b.f[2][idx] = b.f[2];
    s.b.v[0][idx] = b.v[0];
    s.b.v[1][idx] = b.v[1];
    s.b.v[2][idx] = b.v[2];
    s.b.w[idx] = b.w;
}

// Where a neutral of a cell's pick pool lives: a global index of the
// species' arrays, or (REF_LDS set) a slot of the workgroup's birth staging
// (a neutral split off in this call, not yet flushed).  REF_NONE: dropped.
constexpr uint32_t REF_LDS = 0x80000000u;
constexpr uint32_t REF_NONE = 0xFFFFFFFFu;

// species k (0 neutrals, 1 ions, 2 electrons) gains a particle; returns
// where it is staged (see REF_LDS)
__device__ inline uint32_t append(BirthStage& st, int k, Sp& s, uint32_t cell, float fx, float fy, float fz,
                                  const double* v, double w, unsigned long long* overflow) {
    Birth b;
    b.cell = cell;
    b.f[0] = fx;
    b.f[1] = fy;
    b.f[2] = fz;
    b.v[0] = (float)v[0];
    b.v[1] = (float)v[1];
    b.v[2] = (float)v[2];
    b.w = (float)w;
    const uint32_t i = atomicAdd(&st.count[k], 1u);
    if (i < (uint32_t)birth_cap(k)) {
        st.b[birth_off(k) + i] = b;
        return (uint32_t)(birth_off(k) + i) | REF_LDS;
    }
    const uint32_t idx = atomicAdd(&s.d_n[0], 1u);  // staging full
    if (idx >= s.cap) {
        atomicAdd(overflow, 1ull);
        return REF_NONE;
    }
    store_birth(s, idx, b);
    return idx;
}

// the staged births of the workgroup to the species' tails (every thread
// of the workgroup calls this)
__device__ inline void flush_births(BirthStage& st, Sp& N, Sp& I, Sp& E, unsigned long long* overflow) {
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        const uint32_t n = min(st.count[k], (uint32_t)birth_cap(k));
        Sp& s = k == 0 ? N : (k == 1 ? I : E);
        st.base[k] = n ? atomicAdd(&s.d_n[0], n) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; k++) {
        Sp& s = k == 0 ? N : (k == 1 ? I : E);
        const uint32_t n = min(st.count[k], (uint32_t)birth_cap(k));
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
            const uint32_t idx = st.base[k] + i;
            if (idx >= s.cap) {
                atomicAdd(overflow, 1ull);
                continue;
            }
            store_birth(s, idx, st.b[birth_off(k) + i]);
        }
    }
}

// cross-section tables up to this many points are searched in LDS (a
// table's binary search otherwise costs a chain of dependent global loads
// per NTC group)
constexpr int SIG_LDS_MAX = 256;
// A cell's split neutrals of one call that join its pick pool: the first
// POOL_OWN are listed in LDS; later ones in chunks of a global arena
// (SPILL_CHUNK words: the previous chunk's base, then SPILL_CHUNK - 1
// entries), taken with one atomic per chunk.  A split neutral that finds the
// arena full stays out of the pool (counted, reported by pic_mcc_apply).
constexpr int POOL_OWN = 16;
constexpr int SPILL_CHUNK = 32;
constexpr uint32_t SPILL_CAP = 1u << 20;  // words (4 MiB) per MC_MEX object
constexpr uint32_t CHUNK_NONE = 0xFFFFFFFFu;
// The appended part's size, as a fraction of the sorted part, past which it
// is joined with it: indexing costs grow with the appended part, a join
// streams the whole species once, and picks through the index are scattered
// reads.
constexpr double JOIN_FRACTION = 0.75;

template <bool LDS_TABLE>
__global__ void __launch_bounds__(BLOCK)
    k_mcc(GridC g, MccArgs a, Sp N, const uint32_t* __restrict__ offN, const uint32_t* __restrict__ toffN,
          const uint32_t* __restrict__ tidxN, Sp I, Sp E, const uint32_t* __restrict__ offE,
          const uint32_t* __restrict__ toffE, const uint32_t* __restrict__ tidxE, const double* __restrict__ table,
          const double* __restrict__ wsv, unsigned long long* __restrict__ part, uint32_t* __restrict__ spill) {
    // this workgroup's stripe: collisions, ionizations, groups, overflow, step maximum
    unsigned long long* stats = part + (blockIdx.x % NSTRIPE) * SSTRIDE;
    __shared__ double stab[LDS_TABLE ? 2 * SIG_LDS_MAX : 1];
    __shared__ BirthStage births;
    if (threadIdx.x < 3) births.count[threadIdx.x] = 0;
    if (LDS_TABLE) {
        for (int q = threadIdx.x; q < 2 * a.nsig; q += BLOCK) stab[q] = table[q];
        table = stab;
    }
    __syncthreads();
    // this thread's split neutrals in its cell's pool: where they are (a
    // staging slot, or a global index once the staging is full)
    __shared__ uint32_t own[POOL_OWN][BLOCK];
    int c = blockIdx.x * BLOCK + threadIdx.x;
    double wsv_tmp = 0;
    unsigned long long ncoll = 0, niz = 0, ngr = 0, ndead = 0, npool_full = 0;
    // the arena's fill (stripe 0's spare word, cleared with the partials)
    unsigned long long* spill_fill = part + 7;
    uint32_t chunk = CHUNK_NONE;  // this thread's last arena chunk
    if (c < g.ncells) {
        // a cell's electrons: its range of the sorted part, then (toffE) its
        // entries of the appended part's index
        uint32_t e0 = offE[c], e1 = offE[c + 1], te0 = 0, te1 = 0;
        if (toffE) {
            te0 = toffE[c];
            te1 = toffE[c + 1];
        }
        uint32_t n0 = offN[c], n1 = offN[c + 1];
        // a cell's neutrals: its range of the sorted part, then (toffN) its
        // range of the appended part's cell-ordered index tidxN
        uint32_t t0 = 0, t1 = 0;
        if (toffN) {
            t0 = toffN[c];
            t1 = toffN[c + 1];
        }
        // the cell's pick pool (Interactions.cpp:633-635, :698-700): its
        // range of the sorted part (Nb), its entries of the appended part's
        // index (Nt), then the neutrals split off in this call (nown, as the
        // reference appends their indices to the cell's list); removals take
        // the last entry into the removed slot (:719-721)
        int Nb = (int)(n1 - n0), Nt = (int)(t1 - t0), nown = 0;
        const int Neb = (int)(e1 - e0), Ne = Neb + (int)(te1 - te0), Nn0 = Nb + Nt;
        auto loc_of = [&](int slot) -> uint32_t {
            if (slot < Nb) return n0 + (uint32_t)slot;
            if (slot < Nb + Nt) return tidxN[t0 + (uint32_t)(slot - Nb)];
            const int k = slot - Nb - Nt;
            if (k < POOL_OWN) return own[k][threadIdx.x];
            // walk back from the last chunk to the one holding entry k
            const int j = k - POOL_OWN;
            const int last = (nown - POOL_OWN - 1) / (SPILL_CHUNK - 1);
            uint32_t ch = chunk;
            for (int h = last; h > j / (SPILL_CHUNK - 1); h--) ch = spill[ch];
            return spill[ch + 1 + j % (SPILL_CHUNK - 1)];
        };
        if (Ne > 0 && Nn0 > 0) {
            const double wmax = wsv[0];
            // Interactions.cpp:646, the same operation order (an exact tie
            // at .5 then rounds the same way)
            double frac = (double)(Nn0 * Ne) * wmax * a.dt * a.inv_dv;
            int n_groups = (int)(frac + 0.5);
            if (n_groups > Nn0) n_groups = Nn0 - 1;
            const double* tE = table;
            const double* tV = table + a.nsig;
            for (int gi = 0; gi < n_groups; gi++) {
                Draws dr;
                dr.init((uint32_t)c, (uint32_t)gi, a.seed, a.step, TAG_MCC);
                ngr++;
                const int Nn = Nb + Nt + nown;
                int ns = (int)(dr(0) * (double)Nn);
                int es = (int)(dr(1) * (double)Ne);
                const uint32_t pn = loc_of(ns);
                const uint32_t pe = es < Neb ? e0 + (uint32_t)es : tidxE[te0 + (uint32_t)(es - Neb)];
                const bool staged = (pn & REF_LDS) != 0;
                Birth* nb = staged ? &births.b[pn & ~REF_LDS] : nullptr;
                // both particles' weights and velocities requested together
                double wn, vn[3];
                if (staged) {
                    wn = nb->w;
                    vn[0] = nb->v[0];
                    vn[1] = nb->v[1];
                    vn[2] = nb->v[2];
                } else {
                    wn = N.b.w[pn];
                    vn[0] = N.b.v[0][pn];
                    vn[1] = N.b.v[1][pn];
                    vn[2] = N.b.v[2][pn];
                }
                double we = E.b.w[pe];
                double ve[3] = {E.b.v[0][pe], E.b.v[1][pe], E.b.v[2][pe]};
                // (an empty asm that takes them all: otherwise the compiler
                // tests wn first and waits for memory more than once)
                asm volatile("" ::"v"(wn), "v"(we), "v"(vn[0]), "v"(vn[1]), "v"(vn[2]), "v"(ve[0]), "v"(ve[1]),
                             "v"(ve[2]));
                double vr[3] = {vn[0] - ve[0], vn[1] - ve[1], vn[2] - ve[2]};
                double v_rel = sqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
                if (wn == 0.0) continue;  // a weightless neutral (uploaded so) never collides
                double E_rel = a.E_pre_rel * v_rel * v_rel;
                double sigma = sigma_el(tE, tV, a.nsig, E_rel);
                double Wg = wn > we ? wn : we;
                double wsv_g = Wg * sigma * v_rel;
                if (wsv_g > wsv_tmp) wsv_tmp = wsv_g;
                double P = wsv_g / wmax;
                if (!(dr(2) < P)) continue;
                ncoll++;
                // the neutral's cell and offsets (its births' position), read
                // only for accepted groups: random picks from the SoA arrays
                // fetch a whole line per field, and most groups are rejected
                uint32_t ncell;
                float nfx, nfy, nfz;
                if (staged) {
                    ncell = nb->cell;
                    nfx = nb->f[0];
                    nfy = nb->f[1];
                    nfz = nb->f[2];
                } else {
                    ncell = N.b.cell[pn];
                    nfx = N.b.f[0][pn];
                    nfy = N.b.f[1][pn];
                    nfz = N.b.f[2][pn];
                }
                double vnew[3] = {0, 0, 0};
                // equal weights: no state change (only the electron's
                // velocity is written back)
                if (wn != we) {
                    const bool n_heavier = wn > we;
                    if (n_heavier) {  // Interactions.cpp:684-703
                        if (staged)
                            nb->w = (float)(wn - we);
                        else
                            N.b.w[pn] = (float)(wn - we);
                    } else {  // Interactions.cpp:705-726, defined semantics
                        // the electron keeps the neutral's weight; the rest
                        // leaves as an electron of the pre-collision velocity
                        E.b.w[pe] = (float)wn;
                        append(births, 2, E, E.b.cell[pe], E.b.f[0][pe], E.b.f[1][pe], E.b.f[2][pe], ve, we - wn,
                               &stats[3]);
                    }
                    // one call site for both branches (register pressure)
                    const bool iz = collide(a, dr, vn, ve, vnew, sigma);
                    const double wb = n_heavier ? we : wn;  // the weight that collided
                    if (iz) {
                        niz++;
                        append(births, 1, I, ncell, nfx, nfy, nfz, vn, wb, &stats[3]);
                        append(births, 2, E, ncell, nfx, nfy, nfz, vnew, wb, &stats[3]);
                    }
                    if (n_heavier && !iz) {
                        // the split neutral joins the cell's pool (:698-700)
                        const uint32_t r = append(births, 0, N, ncell, nfx, nfy, nfz, vn, we, &stats[3]);
                        if (r != REF_NONE) {
                            if (nown < POOL_OWN) {
                                own[nown][threadIdx.x] = r;
                                nown++;
                            } else {
                                const int j = nown - POOL_OWN;
                                bool room = true;
                                if (j % (SPILL_CHUNK - 1) == 0) {  // a new chunk
                                    const unsigned long long bb =
                                        atomicAdd(spill_fill, (unsigned long long)SPILL_CHUNK);
                                    if (bb + SPILL_CHUNK <= SPILL_CAP) {
                                        spill[bb] = chunk;
                                        chunk = (uint32_t)bb;
                                    } else {
                                        room = false;
                                    }
                                }
                                if (room) {
                                    spill[chunk + 1 + j % (SPILL_CHUNK - 1)] = r;
                                    nown++;
                                } else {
                                    npool_full++;
                                }
                            }
                        }
                    }
                    if (!n_heavier && iz) {
                        // the ionised neutral leaves the cell's pool as in
                        // the reference (Interactions.cpp:719-721): the
                        // pool's last neutral takes its slot and the pool
                        // shrinks by one.  This thread owns the cell's
                        // ranges and its staged splits, so the move is a
                        // plain copy; the vacated entry is removed (DEAD, at
                        // the next whole sort).
                        const uint32_t pl = loc_of(Nn - 1);
                        if (pl != pn) {
                            Birth m;
                            if (pl & REF_LDS) {
                                m = births.b[pl & ~REF_LDS];
                            } else {
                                m.cell = N.b.cell[pl];
#pragma unroll
                                for (int d = 0; d < 3; d++) {
                                    m.f[d] = N.b.f[d][pl];
                                    m.v[d] = N.b.v[d][pl];
                                }
                                m.w = N.b.w[pl];
                            }
                            if (staged)
                                *nb = m;
                            else
                                store_birth(N, pn, m);
                        }
                        if (pl & REF_LDS) {
                            births.b[pl & ~REF_LDS].w = 0.0f;
                            births.b[pl & ~REF_LDS].cell = DEAD;
                        } else {
                            N.b.w[pl] = 0.0f;
                            N.b.cell[pl] = DEAD;
                        }
                        if (nown > 0) {
                            nown--;
                            // the last chunk emptied: back to the previous one
                            if (nown >= POOL_OWN && (nown - POOL_OWN) % (SPILL_CHUNK - 1) == 0)
                                chunk = spill[chunk];
                        } else if (Nt > 0)
                            Nt--;
                        else
                            Nb--;
                        ndead++;
                    }
                }
                E.b.v[0][pe] = (float)ve[0];
                E.b.v[1][pe] = (float)ve[1];
                E.b.v[2][pe] = (float)ve[2];
            }
        }
    }
    if (npool_full) atomicAdd(&stats[6], npool_full);
    if (ndead) atomicAdd(&stats[5], ndead);  // neutrals removed (rare)
    flush_births(births, N, I, E, &stats[3]);
    // block reduction of the counters and the step maximum
    __shared__ double smax[BLOCK / 64];
    __shared__ unsigned long long scnt[BLOCK / 64][3];
    for (int o = 32; o > 0; o >>= 1) {
        wsv_tmp = fmax(wsv_tmp, __shfl_xor(wsv_tmp, o, 64));
        ncoll += __shfl_xor(ncoll, o, 64);
        niz += __shfl_xor(niz, o, 64);
        ngr += __shfl_xor(ngr, o, 64);
    }
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        smax[wv] = wsv_tmp;
        scnt[wv][0] = ncoll;
        scnt[wv][1] = niz;
        scnt[wv][2] = ngr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0;
        unsigned long long t0 = 0, t1 = 0, t2 = 0;
        for (int w = 0; w < BLOCK / 64; w++) {
            m = fmax(m, smax[w]);
            t0 += scnt[w][0];
            t1 += scnt[w][1];
            t2 += scnt[w][2];
        }
        if (t2) {  // non-negative doubles order as their bit patterns
            atomicMax(&stats[4], (unsigned long long)__double_as_longlong(m));
            atomicAdd(&stats[0], t0);
            atomicAdd(&stats[1], t1);
            atomicAdd(&stats[2], t2);
        }
    }
}

// the stripes summed into stats (and the step maximum into wsv[1]);
// W_sigma_v_rel_max <- step maximum when any collision happened
// (Interactions.cpp:751-756).  stats then holds all the host reads back
// (MCC_OUT words: the six counters, wsv[0..1] as bits, the live counts of
// the three species and the neutrals' and electrons' sorted parts), and the
// stripes are cleared for the next call.
constexpr int MCC_OUT = 13;
__global__ void __launch_bounds__(256)
    k_mcc_finish(double* wsv, unsigned long long* part, unsigned long long* stats, uint32_t* caps_n,
                 uint32_t cap_n, uint32_t* caps_i, uint32_t cap_i, uint32_t* caps_e, uint32_t cap_e) {
    if (threadIdx.x == 0) {
        unsigned long long t[6] = {0, 0, 0, 0, 0, 0}, mb = 0;
        for (int s = 0; s < NSTRIPE; s++) {
            for (int q = 0; q < 4; q++) t[q] += part[s * SSTRIDE + q];
            t[4] += part[s * SSTRIDE + 5];
            t[5] += part[s * SSTRIDE + 6];
            mb = part[s * SSTRIDE + 4] > mb ? part[s * SSTRIDE + 4] : mb;
        }
        for (int q = 0; q < 6; q++) stats[q] = t[q];
        wsv[1] = __longlong_as_double((long long)mb);
        if (t[0] > 0) wsv[0] = wsv[1];
        if (caps_n[0] > cap_n) caps_n[0] = cap_n;
        if (caps_i[0] > cap_i) caps_i[0] = cap_i;
        if (caps_e[0] > cap_e) caps_e[0] = cap_e;
        stats[6] = (unsigned long long)__double_as_longlong(wsv[0]);
        stats[7] = (unsigned long long)__double_as_longlong(wsv[1]);
        stats[8] = caps_n[0];
        stats[9] = caps_i[0];
        stats[10] = caps_e[0];
        stats[11] = caps_n[3];
        stats[12] = caps_e[3];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < NSTRIPE * SSTRIDE; q += blockDim.x) part[q] = 0;
}

}  // namespace

extern "C" int pic_mcc_create(pic_ctx* ctx, int sn, int si, int se, const pic_mcc_desc* desc,
                              const double* sig_E, const double* sig_v, int nsig, int* id) {
    if (!ctx || !desc || !sig_E || !sig_v || nsig < 1 || !id) return PIC_ERR_ARG;
    int rc = pic_check_sid(ctx, sn);
    if (!rc) rc = pic_check_sid(ctx, si);
    if (!rc) rc = pic_check_sid(ctx, se);
    if (rc) return rc;
    const Species &N = ctx->sp[sn], &I = ctx->sp[si], &E = ctx->sp[se];
    // constructor guards (Interactions.cpp:479-489)
    if (N.mpw0 < 1e2 * E.mpw0)
        return pic_fail(ctx, PIC_ERR_ARG, "neutrals mpw0 must be >= 100 x electrons mpw0");
    if (!(desc->E_ion_J > 0)) return pic_fail(ctx, PIC_ERR_ARG, "neutral E_ion must be positive");
    if (N.mpw0 < I.mpw0) return pic_fail(ctx, PIC_ERR_ARG, "neutrals mpw0 must be >= ions mpw0");
    for (int q = 1; q < nsig; q++)
        if (!(sig_E[q] > sig_E[q - 1]))
            return pic_fail(ctx, PIC_ERR_ARG, "cross-section energies must be ascending");
    Mcc m;
    m.sn = sn;
    m.si = si;
    m.se = se;
    m.desc = *desc;
    m.nsig = nsig;
    PIC_HIP(ctx, hipMalloc(&m.sig, 2 * (size_t)nsig * sizeof(double)));
    PIC_HIP(ctx, hipMalloc(&m.d_wsv, 2 * sizeof(double)));
    PIC_HIP(ctx, hipMalloc(&m.d_stats, MCC_OUT * sizeof(unsigned long long)));
    PIC_HIP(ctx, hipMalloc(&m.d_part, (size_t)NSTRIPE * SSTRIDE * sizeof(unsigned long long)));
    // (k_mcc_finish clears the stripes after each call)
    PIC_HIP(ctx, hipMemsetAsync(m.d_part, 0, (size_t)NSTRIPE * SSTRIDE * sizeof(unsigned long long),
                                ctx->stream));
    PIC_HIP(ctx, hipMalloc(&m.d_spill, (size_t)SPILL_CAP * sizeof(uint32_t)));
    PIC_HIP(ctx, hipMemcpyAsync(m.sig, sig_E, nsig * sizeof(double), hipMemcpyHostToDevice,
                                ctx->stream));
    PIC_HIP(ctx, hipMemcpyAsync(m.sig + nsig, sig_v, nsig * sizeof(double), hipMemcpyHostToDevice,
                                ctx->stream));
    // W_sigma_v_rel_max = 1e-14 * max(mpw0) unless given (Interactions.cpp:129, :534)
    double w0 = desc->wsv_max_init > 0 ? desc->wsv_max_init : 1e-14 * std::max(E.mpw0, N.mpw0);
    double hw[2] = {w0, 0.0};
    PIC_HIP(ctx, hipMemcpyAsync(m.d_wsv, hw, sizeof(hw), hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    m.wsv_host = w0;
    ctx->mcc.push_back(m);
    *id = (int)ctx->mcc.size() - 1;
    return PIC_OK;
}

extern "C" int pic_mcc_apply(pic_ctx* ctx, int id, double dt, uint32_t seed, uint32_t step,
                             pic_mcc_stats* stats) {
    if (!ctx || id < 0 || id >= (int)ctx->mcc.size()) return PIC_ERR_ARG;
    Mcc& m = ctx->mcc[id];
    Species &N = ctx->sp[m.sn], &I = ctx->sp[m.si], &E = ctx->sp[m.se];
    int rc = 0;
    // the neutrals (they move once per many steps): the births appended
    // since their last whole sort are indexed by cell, and joined with the
    // sorted part once they pass JOIN_FRACTION of it; after anything else a
    // whole counting sort
    if (!N.sorted) {
        if (N.tail_ok && N.n_upper >= N.bulk_n) {
            rc = pic_sort_tail(ctx, m.sn);
            if (!rc && (double)(N.n_upper - N.bulk_n) > JOIN_FRACTION * (double)N.bulk_n)
                rc = pic_join_tail(ctx, m.sn);
        } else {
            rc = pic_sort(ctx, m.sn);
        }
    }
    // the electrons likewise (their births are appended; none is removed)
    if (!rc && !E.sorted) {
        if (E.tail_ok && E.n_upper >= E.bulk_n) {
            rc = pic_sort_tail(ctx, m.se);
            if (!rc && (double)(E.n_upper - E.bulk_n) > JOIN_FRACTION * (double)E.bulk_n)
                rc = pic_join_tail(ctx, m.se);
        } else {
            rc = pic_sort(ctx, m.se);
        }
    }
    if (rc) return rc;
    const bool two = !N.sorted && N.two_part;
    const bool twoE = !E.sorted && E.two_part;
    const GridC& g = ctx->g;
    MccArgs a;
    a.m_n = N.mass;
    a.m_e = E.mass;
    a.sum_m = N.mass + E.mass;
    a.m_red = N.mass * E.mass / (N.mass + E.mass);
    a.E_pre_rel = 0.5 * a.m_red / QE;
    a.E_pre_ele = ME * 0.5 / QE;
    a.two_q_me = 2 * QE / ME;
    a.E_ion_eV = m.desc.E_ion_J / QE;
    a.c0 = m.desc.c0;
    a.c1 = m.desc.c1;
    a.c2 = m.desc.c2;
    a.B_inc = m.desc.B_inc;
    a.dt = dt;
    a.inv_dv = 1 / (g.dx[0] * g.dx[1] * g.dx[2]);  // World::getCellVolume (World.cpp:101-103)
    a.seed = seed;
    a.step = step;
    a.nsig = m.nsig;
    Sp sn{N.buf[N.cur], N.d_n, (uint32_t)N.cap};
    Sp si{I.buf[I.cur], I.d_n, (uint32_t)I.cap};
    Sp se{E.buf[E.cur], E.d_n, (uint32_t)E.cap};
    {
        PicProfScope ps(ctx, PIC_K_MCC);
        if (m.nsig <= SIG_LDS_MAX)
            k_mcc<true><<<grid_for((size_t)g.ncells, BLOCK), BLOCK, 0, ctx->stream>>>(
                g, a, sn, N.off, two ? N.toff : nullptr, N.tidx, si, se, E.off, twoE ? E.toff : nullptr, E.tidx,
                m.sig, m.d_wsv,
                m.d_part, m.d_spill);
        else
            k_mcc<false><<<grid_for((size_t)g.ncells, BLOCK), BLOCK, 0, ctx->stream>>>(
                g, a, sn, N.off, two ? N.toff : nullptr, N.tidx, si, se, E.off, twoE ? E.toff : nullptr, E.tidx,
                m.sig, m.d_wsv,
                m.d_part, m.d_spill);
        k_mcc_finish<<<1, 256, 0, ctx->stream>>>(m.d_wsv, m.d_part, m.d_stats, N.d_n, (uint32_t)N.cap,
                                                 I.d_n, (uint32_t)I.cap, E.d_n, (uint32_t)E.cap);
    }
    PIC_HIP(ctx, hipGetLastError());
    // the step's results through pinned memory (one asynchronous copy), one wait
    rc = pic_ensure_pin(ctx, MCC_OUT * sizeof(unsigned long long));
    if (rc) return rc;
    unsigned long long* hs = reinterpret_cast<unsigned long long*>(ctx->pin);  // MCC_OUT words
    PIC_HIP(ctx, hipMemcpyAsync(hs, m.d_stats, MCC_OUT * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                ctx->stream));
    PIC_HIP(ctx, pic_spin_wait(ctx->stream));
    double hw[2];
    std::memcpy(&hw[0], &hs[6], sizeof(double));
    std::memcpy(&hw[1], &hs[7], sizeof(double));
    N.n_upper = (size_t)hs[8];
    I.n_upper = (size_t)hs[9];
    E.n_upper = (size_t)hs[10];
    N.bulk_n = (size_t)hs[11];  // the sorted parts (the live counts of the last whole sorts)
    E.bulk_n = (size_t)hs[12];
    m.wsv_host = hw[0];
    if (hs[0] || hs[1]) {
        N.sorted = I.sorted = E.sorted = false;
        N.hist_valid = I.hist_valid = E.hist_valid = false;
    }
    // births only appended: the neutrals' and electrons' next sorts may
    // take just those, unless a neutral was removed (a removed particle
    // leaves the species only in a whole sort); the ions are sorted whole
    if (hs[4]) N.tail_ok = false;
    I.tail_ok = false;
    I.two_part = false;
    if (stats) {
        stats->collisions = (long long)hs[0];
        stats->ionizations = (long long)hs[1];
        stats->groups = (long long)hs[2];
        stats->wsv_max = hw[0];
        stats->pool_overflow = (long long)hs[5];
    }
    if (hs[3]) return pic_fail(ctx, PIC_ERR_CAPACITY, "species capacity exceeded by DSMC births");
    return PIC_OK;
}
