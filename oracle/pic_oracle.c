/*
 * pic_oracle.c -- FP64 CPU oracle for the PIC/DSMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see pic_oracle.h).  A clean-room restatement of
 * the reference algorithms; operation order follows the reference wherever it
 * decides the last bits of an FP64 result, so that fixtures written by the
 * reference itself can be compared bit-for-bit where that is meaningful.
 * All file:line citations are relative to the reference repository root;
 * "v3/" abbreviates ch4/v3/src/.
 */
#include "pic_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define IDX(g, i, j, k) ((size_t)((k) * (g)->nj + (j)) * (size_t)(g)->ni + (size_t)(i))

/* ------------------------------------------------------------------------ */
/* geometry                                                                  */
/* ------------------------------------------------------------------------ */

/* v3/World.cpp:63-78 */
void orc_grid_init(orc_grid* g, int ni, int nj, int nk, const double* x0, const double* xm) {
    int nn[3] = {ni, nj, nk};
    g->ni = ni;
    g->nj = nj;
    g->nk = nk;
    for (int d = 0; d < 3; d++) {
        g->x0[d] = x0[d];
        g->xm[d] = xm[d];
        g->dx[d] = (xm[d] - x0[d]) / (nn[d] - 1);
        g->inv_dx[d] = 1 / g->dx[d];
    }
}

/* v3/World.cpp:353-367 */
void orc_node_volumes(const orc_grid* g, double* vol) {
    double base = g->dx[0] * g->dx[1] * g->dx[2];
    for (int k = 0; k < g->nk; k++)
        for (int j = 0; j < g->nj; j++)
            for (int i = 0; i < g->ni; i++) {
                double v = base;
                if (i == 0 || i == g->ni - 1) v *= 0.5;
                if (j == 0 || j == g->nj - 1) v *= 0.5;
                if (k == 0 || k == g->nk - 1) v *= 0.5;
                vol[IDX(g, i, j, k)] = v;
            }
}

/* v3/World.cpp:201-205 */
int orc_in_bounds(const orc_grid* g, const double* pos) {
    for (int d = 0; d < 3; d++)
        if (pos[d] < g->x0[d] || pos[d] >= g->xm[d]) return 0;
    return 1;
}

/* v3/World.cpp:293-301 + v3/Object.cpp:231-238 (inclusive box test) */
int orc_in_object(const orc_rect* r, int nrect, const double* pos) {
    for (int o = 0; o < nrect; o++) {
        int inside = 1;
        for (int d = 0; d < 3; d++)
            if (fabs(pos[d] - r[o].c[d]) > r[o].half[d]) inside = 0;
        if (inside) return o + 1;
    }
    return 0;
}

/* v3/World.cpp:276-292 (object_id is the bool result, so always 1) */
void orc_object_id(const orc_grid* g, const orc_rect* r, int nrect, int32_t* object_id,
                   int32_t* node_type, double* phi) {
    for (int o = 0; o < nrect; o++)
        for (int k = 0; k < g->nk; k++)
            for (int j = 0; j < g->nj; j++)
                for (int i = 0; i < g->ni; i++) {
                    double x[3] = {g->x0[0] + i * g->dx[0], g->x0[1] + j * g->dx[1],
                                   g->x0[2] + k * g->dx[2]};
                    if (orc_in_object(&r[o], 1, x)) {
                        size_t u = IDX(g, i, j, k);
                        object_id[u] = 1;
                        node_type[u] = 2;
                        phi[u] = r[o].phi;
                    }
                }
}

/* v3/World.cpp:263-273 */
void orc_dirichlet_walls(const orc_grid* g, int32_t* node_type) {
    for (int k = 0; k < g->nk; k++)
        for (int j = 0; j < g->nj; j++)
            for (int i = 0; i < g->ni; i++)
                if (i == 0 || i == g->ni - 1 || j == 0 || j == g->nj - 1 || k == 0 ||
                    k == g->nk - 1)
                    node_type[IDX(g, i, j, k)] = 2;
}

/* v3/World.cpp:123-127 */
static inline void xtol(const orc_grid* g, const double* x, double* lc) {
    for (int d = 0; d < 3; d++) lc[d] = (x[d] - g->x0[d]) * g->inv_dx[d];
}

/* ------------------------------------------------------------------------ */
/* CIC weighting                                                            */
/* ------------------------------------------------------------------------ */

typedef struct {
    size_t u000, u001, u010, u011, u100, u101, u110, u111;
    double a, b, c, d, dk, one_dk; /* a=(1-di)(1-dj) b=(1-di)dj c=di(1-dj) d=di dj */
} cic_t;

/* v3/Field.h:201-232 weight construction (same product order) */
static inline void cic_weights(const orc_grid* g, const double* lc, cic_t* w) {
    int i = (int)lc[0], j = (int)lc[1], k = (int)lc[2];
    double di = lc[0] - i, dj = lc[1] - j, dk = lc[2] - k;
    double one_dj = 1 - dj, one_di = 1 - di, one_dk = 1 - dk;
    w->a = one_di * one_dj;
    w->b = one_di * dj;
    w->c = di * one_dj;
    w->d = di * dj;
    w->dk = dk;
    w->one_dk = one_dk;
    w->u000 = IDX(g, i, j, k);
    w->u001 = IDX(g, i, j, k + 1);
    w->u010 = IDX(g, i, j + 1, k);
    w->u011 = IDX(g, i, j + 1, k + 1);
    w->u100 = IDX(g, i + 1, j, k);
    w->u101 = IDX(g, i + 1, j, k + 1);
    w->u110 = IDX(g, i + 1, j + 1, k);
    w->u111 = IDX(g, i + 1, j + 1, k + 1);
}

double orc_gather1(const orc_grid* g, const double* f, const double* lc) {
    cic_t w;
    cic_weights(g, lc, &w);
    return f[w.u000] * w.a * w.one_dk + f[w.u001] * w.a * w.dk + f[w.u010] * w.b * w.one_dk +
           f[w.u011] * w.b * w.dk + f[w.u100] * w.c * w.one_dk + f[w.u101] * w.c * w.dk +
           f[w.u110] * w.d * w.one_dk + f[w.u111] * w.d * w.dk;
}

void orc_gather3(const orc_grid* g, const double* f, const double* lc, double* out) {
    cic_t w;
    cic_weights(g, lc, &w);
    for (int c = 0; c < 3; c++)
        out[c] = f[3 * w.u000 + c] * w.a * w.one_dk + f[3 * w.u001 + c] * w.a * w.dk +
                 f[3 * w.u010 + c] * w.b * w.one_dk + f[3 * w.u011 + c] * w.b * w.dk +
                 f[3 * w.u100 + c] * w.c * w.one_dk + f[3 * w.u101 + c] * w.c * w.dk +
                 f[3 * w.u110 + c] * w.d * w.one_dk + f[3 * w.u111 + c] * w.d * w.dk;
}

/* v3/Field.h:157-199: the value is folded into the (i,j) weights first */
void orc_scatter1(const orc_grid* g, double* f, const double* lc, double val) {
    int i = (int)lc[0], j = (int)lc[1], k = (int)lc[2];
    double di = lc[0] - i, dj = lc[1] - j, dk = lc[2] - k;
    double one_di = 1 - di, one_dj = 1 - dj, one_dk = 1 - dk;
    double v00 = val * one_di * one_dj;
    double v01 = val * one_di * dj;
    double v10 = val * di * one_dj;
    double v11 = val * di * dj;
    f[IDX(g, i, j, k)] += v00 * one_dk;
    f[IDX(g, i, j, k + 1)] += v00 * dk;
    f[IDX(g, i, j + 1, k)] += v01 * one_dk;
    f[IDX(g, i, j + 1, k + 1)] += v01 * dk;
    f[IDX(g, i + 1, j, k)] += v10 * one_dk;
    f[IDX(g, i + 1, j, k + 1)] += v10 * dk;
    f[IDX(g, i + 1, j + 1, k)] += v11 * one_dk;
    f[IDX(g, i + 1, j + 1, k + 1)] += v11 * dk;
}

static void scatter3(const orc_grid* g, double* f, const double* lc, const double* val) {
    int i = (int)lc[0], j = (int)lc[1], k = (int)lc[2];
    double di = lc[0] - i, dj = lc[1] - j, dk = lc[2] - k;
    double one_di = 1 - di, one_dj = 1 - dj, one_dk = 1 - dk;
    for (int c = 0; c < 3; c++) {
        double v00 = val[c] * one_di * one_dj;
        double v01 = val[c] * one_di * dj;
        double v10 = val[c] * di * one_dj;
        double v11 = val[c] * di * dj;
        f[3 * IDX(g, i, j, k) + c] += v00 * one_dk;
        f[3 * IDX(g, i, j, k + 1) + c] += v00 * dk;
        f[3 * IDX(g, i, j + 1, k) + c] += v01 * one_dk;
        f[3 * IDX(g, i, j + 1, k + 1) + c] += v01 * dk;
        f[3 * IDX(g, i + 1, j, k) + c] += v10 * one_dk;
        f[3 * IDX(g, i + 1, j, k + 1) + c] += v10 * dk;
        f[3 * IDX(g, i + 1, j + 1, k) + c] += v11 * one_dk;
        f[3 * IDX(g, i + 1, j + 1, k + 1) + c] += v11 * dk;
    }
}

/* ------------------------------------------------------------------------ */
/* particles                                                                 */
/* ------------------------------------------------------------------------ */

/* one particle of v3/Species.cpp:369-388: kick with E at the old position,
 * drift, then inObject/inBounds.  Returns 1 if the particle survives. */
static inline int push_one(const orc_grid* g, const double* ef, double* p, double qmdt, double dt,
                           const orc_rect* rects, int nrect) {
    double lc[3], e[3];
    xtol(g, p, lc);
    orc_gather3(g, ef, lc, e);
    for (int d = 0; d < 3; d++) p[3 + d] += e[d] * qmdt;
    for (int d = 0; d < 3; d++) p[d] += p[3 + d] * dt;
    int in_obj = orc_in_object(rects, nrect, p);
    if (!orc_in_bounds(g, p)) return 0;
    if (in_obj) return 0;
    return 1;
}

/* v3/Species.cpp:356-399 */
size_t orc_push(const orc_grid* g, const double* ef, double* parts, size_t n, double charge,
                double mass, double dt, const orc_rect* rects, int nrect, int stable) {
    double qmdt = dt * charge / mass; /* v3/Species.cpp:372 */
    if (stable) {
        size_t out = 0;
        for (size_t p = 0; p < n; p++) {
            double* q = parts + 7 * p;
            if (push_one(g, ef, q, qmdt, dt, rects, nrect)) {
                if (out != p) memmove(parts + 7 * out, q, 7 * sizeof(double));
                out++;
            }
        }
        return out;
    }
    size_t np = n;
    for (size_t p = 0; p < np; p++) {
        double* q = parts + 7 * p;
        if (!push_one(g, ef, q, qmdt, dt, rects, nrect)) {
            memcpy(q, parts + 7 * (np - 1), 7 * sizeof(double));
            np--;
            p--;
        }
    }
    return np;
}

/* v3/Species.cpp:401-416 + v3/Field.h:563-583 */
void orc_number_density(const orc_grid* g, const double* parts, size_t n, const double* vol,
                        double* den) {
    size_t nn = (size_t)g->ni * g->nj * g->nk;
    memset(den, 0, nn * sizeof(double));
    for (size_t p = 0; p < n; p++) {
        double lc[3];
        xtol(g, parts + 7 * p, lc);
        orc_scatter1(g, den, lc, parts[7 * p + 6]);
    }
    for (size_t u = 0; u < nn; u++) den[u] = vol[u] != 0 ? den[u] / vol[u] : 0;
}

/* v3/World.cpp:193-200 */
void orc_charge_accumulate(size_t nn, double* rho, double charge, const double* den) {
    if (charge == 0) return;
    for (size_t u = 0; u < nn; u++) rho[u] += charge * den[u];
}

/* v3/Species.cpp:767-776 */
void orc_sample_moments(const orc_grid* g, const double* parts, size_t n, double* n_sum,
                        double* nv_sum, double* nuu, double* nvv, double* nww) {
    for (size_t p = 0; p < n; p++) {
        const double* q = parts + 7 * p;
        double lc[3], wv[3];
        double w = q[6];
        xtol(g, q, lc);
        orc_scatter1(g, n_sum, lc, w);
        for (int d = 0; d < 3; d++) wv[d] = q[3 + d] * w;
        scatter3(g, nv_sum, lc, wv);
        orc_scatter1(g, nuu, lc, w * q[3] * q[3]);
        orc_scatter1(g, nvv, lc, w * q[4] * q[4]);
        orc_scatter1(g, nww, lc, w * q[5] * q[5]);
    }
}

/* v3/Species.cpp:777-804 and v3/Field.h:586-605 (vector / scalar field) */
void orc_gas_properties(const orc_grid* g, double mass, const double* n_sum, const double* nv_sum,
                        const double* nuu, const double* nvv, const double* nww, double* vel,
                        double* T) {
    size_t nn = (size_t)g->ni * g->nj * g->nk;
    for (size_t u = 0; u < nn; u++) {
        double cnt = n_sum[u];
        if (cnt != 0) {
            double inv = 1.0 / cnt;
            for (int d = 0; d < 3; d++) vel[3 * u + d] = nv_sum[3 * u + d] * inv;
        } else {
            for (int d = 0; d < 3; d++) vel[3 * u + d] = 0;
        }
        if (cnt <= 0) {
            T[u] = 0;
            continue;
        }
        double uu = nuu[u] / cnt - vel[3 * u] * vel[3 * u];
        double vv = nvv[u] / cnt - vel[3 * u + 1] * vel[3 * u + 1];
        double ww = nww[u] / cnt - vel[3 * u + 2] * vel[3 * u + 2];
        T[u] = mass / (2 * ORC_KB) * (uu + vv + ww);
    }
}

/* v3/Species.cpp:813-819 + v3/World.cpp:130-133 */
void orc_macro_count(const orc_grid* g, const double* parts, size_t n, double* count) {
    int ci = g->ni - 1, cj = g->nj - 1, ck = g->nk - 1;
    memset(count, 0, (size_t)ci * cj * ck * sizeof(double));
    for (size_t p = 0; p < n; p++) {
        double lc[3];
        xtol(g, parts + 7 * p, lc);
        int i = (int)lc[0], j = (int)lc[1], k = (int)lc[2];
        count[((size_t)k * cj + j) * ci + i] += 1;
    }
}

static inline int cell_of(const orc_grid* g, const double* x) {
    double lc[3];
    xtol(g, x, lc);
    return ((int)lc[2] * (g->nj - 1) + (int)lc[1]) * (g->ni - 1) + (int)lc[0];
}

/* v3/Species.cpp:905-915 as CSR: lists keep particle order */
void orc_sort_indexes(const orc_grid* g, const double* parts, size_t n, int32_t* offsets,
                      int32_t* index) {
    int nc = (g->ni - 1) * (g->nj - 1) * (g->nk - 1);
    memset(offsets, 0, (size_t)(nc + 1) * sizeof(int32_t));
    for (size_t p = 0; p < n; p++) offsets[cell_of(g, parts + 7 * p) + 1]++;
    for (int c = 0; c < nc; c++) offsets[c + 1] += offsets[c];
    int32_t* cur = (int32_t*)malloc((size_t)nc * sizeof(int32_t));
    memcpy(cur, offsets, (size_t)nc * sizeof(int32_t));
    for (size_t p = 0; p < n; p++) index[cur[cell_of(g, parts + 7 * p)]++] = (int32_t)p;
    free(cur);
}

/* v3/Species.cpp:732-763 */
void orc_species_sums(const orc_grid* g, const double* parts, size_t n, double mass, double charge,
                      const double* phi, double* out) {
    double micro = 0, mom[3] = {0, 0, 0}, ke = 0, pe = 0;
    for (size_t p = 0; p < n; p++) {
        const double* q = parts + 7 * p;
        micro += q[6];
        for (int d = 0; d < 3; d++) mom[d] += q[3 + d] * q[6];
        ke += q[6] * (q[3] * q[3] + q[4] * q[4] + q[5] * q[5]);
        if (phi) {
            double lc[3];
            xtol(g, q, lc);
            pe += q[6] * orc_gather1(g, phi, lc);
        }
    }
    out[0] = micro;
    for (int d = 0; d < 3; d++) out[1 + d] = mom[d] * mass;
    out[4] = ke * mass * 0.5;
    out[5] = pe * charge;
}

/* v3/World.cpp:108-118 (i outer, k inner summation order) */
double orc_field_energy(const orc_grid* g, const double* ef, const double* vol) {
    double pe = 0;
    for (int i = 0; i < g->ni; i++)
        for (int j = 0; j < g->nj; j++)
            for (int k = 0; k < g->nk; k++) {
                size_t u = IDX(g, i, j, k);
                const double* e = ef + 3 * u;
                pe += (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) * vol[u];
            }
    return 0.5 * ORC_EPS0 * pe;
}

/* ------------------------------------------------------------------------ */
/* electric field, v3/PotentialSolver.cpp:354-408                           */
/* ------------------------------------------------------------------------ */
void orc_efield(const orc_grid* g, const double* phi, double* ef) {
    double inv2[3];
    for (int d = 0; d < 3; d++) inv2[d] = 1.0 / (2 * g->dx[d]);
    int n[3] = {g->ni, g->nj, g->nk};
    size_t st[3] = {1, (size_t)g->ni, (size_t)g->ni * g->nj};
    for (int k = 0; k < g->nk; k++)
        for (int j = 0; j < g->nj; j++)
            for (int i = 0; i < g->ni; i++) {
                int ijk[3] = {i, j, k};
                size_t u = IDX(g, i, j, k);
                for (int d = 0; d < 3; d++) {
                    size_t s = st[d];
                    double v;
                    if (ijk[d] == 0)
                        v = (3 * phi[u] - 4 * phi[u + s] + phi[u + 2 * s]) * inv2[d];
                    else if (ijk[d] == n[d] - 1)
                        v = (-phi[u - 2 * s] + 4 * phi[u - s] - 3 * phi[u]) * inv2[d];
                    else
                        v = (phi[u - s] - phi[u + s]) * inv2[d];
                    ef[3 * u + d] = v;
                }
            }
}

/* ------------------------------------------------------------------------ */
/* Gauss-Seidel / SOR                                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    double d2x, d2y, d2z, ieps, twos, itwos;
} stencil_t;
static void stencil_init(const orc_grid* g, stencil_t* s) {
    s->d2x = 1.0 / (g->dx[0] * g->dx[0]);
    s->d2y = 1.0 / (g->dx[1] * g->dx[1]);
    s->d2z = 1.0 / (g->dx[2] * g->dx[2]);
    s->ieps = 1.0 / ORC_EPS0;
    s->twos = 2.0 * (s->d2x + s->d2y + s->d2z);
    s->itwos = 1.0 / s->twos;
}

/* ch2/v1/PotentialSolver.cpp:8-87 (sweep order i, j, k with k fastest) */
int orc_sor_dirichlet(const orc_grid* g, double* phi, const double* rho, int max_it, double tol,
                      double omega, int* iters, double* l2) {
    stencil_t s;
    stencil_init(g, &s);
    size_t sj = g->ni, sk = (size_t)g->ni * g->nj;
    double L2 = 0;
    int it;
    int converged = 0;
    for (it = 0; it < max_it; it++) {
        for (int i = 1; i < g->ni - 1; i++)
            for (int j = 1; j < g->nj - 1; j++)
                for (int k = 1; k < g->nk - 1; k++) {
                    size_t u = IDX(g, i, j, k);
                    double nw = (rho[u] * s.ieps + (phi[u - 1] + phi[u + 1]) * s.d2x +
                                 (phi[u - sj] + phi[u + sj]) * s.d2y +
                                 (phi[u - sk] + phi[u + sk]) * s.d2z) *
                                s.itwos;
                    phi[u] = phi[u] + omega * (nw - phi[u]);
                }
        if (it % 25 == 0) {
            double sum = 0;
            for (int i = 1; i < g->ni - 1; i++)
                for (int j = 1; j < g->nj - 1; j++)
                    for (int k = 1; k < g->nk - 1; k++) {
                        size_t u = IDX(g, i, j, k);
                        double R = -phi[u] * s.twos + rho[u] * s.ieps +
                                   (phi[u - 1] + phi[u + 1]) * s.d2x +
                                   (phi[u - sj] + phi[u + sj]) * s.d2y +
                                   (phi[u - sk] + phi[u + sk]) * s.d2z;
                        sum += R * R;
                    }
            L2 = sqrt(sum / ((double)g->ni * g->nj * g->nk));
            if (L2 < tol) {
                converged = 1;
                break;
            }
        }
    }
    if (iters) *iters = converged ? it + 1 : it;
    if (l2) *l2 = L2;
    return converged;
}

/* v3/PotentialSolver.cpp:69-166 */
int orc_gs_v3(const orc_grid* g, const int32_t* object_id, double* phi, const double* rho,
              double phi0, double n0, double Te0, int max_it, double tol, double omega, int* iters,
              double* l2) {
    stencil_t s;
    stencil_init(g, &s);
    size_t sj = g->ni, sk = (size_t)g->ni * g->nj;
    int ni = g->ni, nj = g->nj, nk = g->nk;
    double L2 = 0;
    int it, converged = 0;
    for (it = 0; it < max_it; it++) {
        for (int i = 0; i < ni; i++)
            for (int j = 0; j < nj; j++)
                for (int k = 0; k < nk; k++) {
                    size_t u = IDX(g, i, j, k);
                    if (object_id[u] > 0) continue;
                    if (i == 0)
                        phi[u] = phi[u + 1];
                    else if (i == ni - 1)
                        phi[u] = phi[u - 1];
                    else if (j == 0)
                        phi[u] = phi[u + sj];
                    else if (j == nj - 1)
                        phi[u] = phi[u - sj];
                    else if (k == 0)
                        phi[u] = phi[u + sk];
                    else if (k == nk - 1)
                        phi[u] = phi[u - sk];
                    else {
                        double ne = n0 * exp((phi[u] - phi0) / Te0);
                        double nw = ((rho[u] - ORC_QE * ne) * s.ieps +
                                     (phi[u - 1] + phi[u + 1]) * s.d2x +
                                     (phi[u - sj] + phi[u + sj]) * s.d2y +
                                     (phi[u - sk] + phi[u + sk]) * s.d2z) *
                                    s.itwos;
                        phi[u] = phi[u] + omega * (nw - phi[u]);
                    }
                }
        if (it % 25 == 0) {
            double sum = 0;
            for (int i = 0; i < ni; i++)
                for (int j = 0; j < nj; j++)
                    for (int k = 0; k < nk; k++) {
                        size_t u = IDX(g, i, j, k);
                        double R;
                        if (object_id[u] > 0) continue;
                        if (i == 0)
                            R = phi[u] - phi[u + 1];
                        else if (i == ni - 1)
                            R = phi[u] - phi[u - 1];
                        else if (j == 0)
                            R = phi[u] - phi[u + sj];
                        else if (j == nj - 1)
                            R = phi[u] - phi[u - sj];
                        else if (k == 0)
                            R = phi[u] - phi[u + sk];
                        else if (k == nk - 1)
                            R = phi[u] - phi[u - sk];
                        else {
                            double ne = n0 * exp((phi[u] - phi0) / Te0);
                            R = -phi[u] * s.twos + (rho[u] - ORC_QE * ne) * s.ieps +
                                (phi[u - 1] + phi[u + 1]) * s.d2x +
                                (phi[u - sj] + phi[u + sj]) * s.d2y +
                                (phi[u - sk] + phi[u + sk]) * s.d2z;
                        }
                        sum += R * R;
                    }
            L2 = sqrt(sum / ((double)ni * nj * nk));
            if (L2 < tol) {
                converged = 1;
                break;
            }
        }
    }
    if (iters) *iters = converged ? it + 1 : it;
    if (l2) *l2 = L2;
    return converged;
}

/* ------------------------------------------------------------------------ */
/* assembled 7-entry matrix (v3/Matrix.h:12-62, v3/Matrix.cpp:99-284)        */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    int32_t* col; /* 7 per row, -1 terminated */
    double* a;
} mat7;

static void mat_alloc(mat7* m, int n) {
    m->n = n;
    m->col = (int32_t*)malloc((size_t)n * 7 * sizeof(int32_t));
    m->a = (double*)calloc((size_t)n * 7, sizeof(double));
    for (size_t i = 0; i < (size_t)n * 7; i++) m->col[i] = -1;
}
static void mat_free(mat7* m) {
    free(m->col);
    free(m->a);
}
/* Matrix::operator()(r,c): find or append the column (Matrix.cpp:180-192) */
static double* mat_at(mat7* m, int r, int c) {
    int32_t* col = m->col + 7 * (size_t)r;
    int i;
    for (i = 0; i < 7; i++) {
        if (col[i] == c) break;
        if (col[i] < 0) {
            col[i] = c;
            break;
        }
    }
    return m->a + 7 * (size_t)r + i;
}
/* Matrix::operator* (Matrix.cpp:117-141): entries in insertion order */
static void mat_vec(const mat7* m, const double* v, double* out) {
    for (int u = 0; u < m->n; u++) {
        double acc = 0;
        const int32_t* col = m->col + 7 * (size_t)u;
        const double* a = m->a + 7 * (size_t)u;
        for (int i = 0; i < 7; i++) {
            if (col[i] < 0) break;
            acc += a[i] * v[col[i]];
        }
        out[u] = acc;
    }
}
static double vdot(const double* a, const double* b, int n) {
    double s = 0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}
static double vnorm(const double* a, int n) { return sqrt(vdot(a, a, n) / n); }

/* v3/PotentialSolver.cpp:243-313 */
static int pcg_linear(const mat7* A, const double* Mdiag, double* x, const double* b, int max_it,
                      double tol, int* iters_out, double* l2_out) {
    int n = A->n;
    double* g = (double*)malloc(n * sizeof(double));
    double* s = (double*)malloc(n * sizeof(double));
    double* d = (double*)malloc(n * sizeof(double));
    double* z = (double*)malloc(n * sizeof(double));
    mat_vec(A, x, g);
    for (int i = 0; i < n; i++) g[i] = g[i] - b[i];
    for (int i = 0; i < n; i++) s[i] = 0 + Mdiag[i] * g[i];
    for (int i = 0; i < n; i++) d[i] = s[i] * -1.0;
    int converged = 0, it;
    double L2 = 0;
    int ret = 0;
    for (it = 0; it < max_it; it++) {
        mat_vec(A, d, z);
        double alpha = vdot(g, s, n);
        double beta = vdot(d, z, n);
        if (beta == 0) {
            ret = -1;
            break;
        }
        double r = alpha / beta;
        for (int i = 0; i < n; i++) x[i] = x[i] + d[i] * r;
        for (int i = 0; i < n; i++) g[i] = g[i] + z[i] * r;
        for (int i = 0; i < n; i++) s[i] = 0 + Mdiag[i] * g[i];
        beta = alpha;
        alpha = vdot(g, s, n);
        double r2 = alpha / beta;
        for (int i = 0; i < n; i++) d[i] = d[i] * r2 - s[i];
        L2 = vnorm(g, n);
        if (L2 < tol) {
            converged = 1;
            break;
        }
    }
    if (iters_out) *iters_out = converged ? it + 1 : it;
    if (l2_out) *l2_out = L2;
    free(g);
    free(s);
    free(d);
    free(z);
    if (ret < 0) return -1;
    return converged;
}

/* v3/PotentialSolver.cpp:315-347 */
static int gs_linear(const mat7* A, double* x, const double* b, int max_it, double tol) {
    int n = A->n;
    double* R = (double*)malloc(n * sizeof(double));
    int converged = 0;
    for (int it = 0; it < max_it; it++) {
        for (int i = 0; i < n; i++) {
            const int32_t* col = A->col + 7 * (size_t)i;
            const double* a = A->a + 7 * (size_t)i;
            double sum = 0, aii = 0;
            for (int e = 0; e < 7; e++) {
                if (col[e] < 0) break;
                sum += a[e] * x[col[e]];
                if (col[e] == i) aii = a[e];
            }
            double S = sum - aii * x[i];
            double nw = (b[i] - S) / aii;
            x[i] = x[i] + 1.0 * (nw - x[i]);
        }
        if (it % 25 == 0) {
            mat_vec(A, x, R);
            for (int i = 0; i < n; i++) R[i] = R[i] - b[i];
            if (vnorm(R, n) < tol) {
                converged = 1;
                break;
            }
        }
    }
    free(R);
    return converged;
}

/* v3/PotentialSolver.cpp:178-241 with the matrix of :414-471 */
int orc_nrpcg(const orc_grid* g, const int32_t* object_id, const int32_t* node_type_in,
              double* phi, const double* rho, double phi0, double n0, double Te0, int pcg_max_it,
              double tol, int axis_swap, int* newton_iters, int* pcg_iters_total,
              double* last_l2) {
    enum { REG = 0, NEU = 1, DIR = 2 };
    int ni = g->ni, nj = g->nj, nk = g->nk;
    int n = ni * nj * nk;
    stencil_t s;
    stencil_init(g, &s);
    double inv_d[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
    /* reference placement: +-ni*nj gets inv_d2x, +-1 gets inv_d2z */
    double cx = axis_swap ? s.d2z : s.d2x;
    double cz = axis_swap ? s.d2x : s.d2z;
    mat7 A;
    mat_alloc(&A, n);
    int* nt = (int*)malloc(n * sizeof(int));
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                int u = (k * nj + j) * ni + i;
                if (object_id[u] > 0 || node_type_in[u] == DIR) {
                    *mat_at(&A, u, u) = 1;
                    nt[u] = DIR;
                    continue;
                }
                nt[u] = NEU;
                if (i == 0) {
                    *mat_at(&A, u, u) = inv_d[0];
                    *mat_at(&A, u, u + 1) = -inv_d[0];
                } else if (i == ni - 1) {
                    *mat_at(&A, u, u) = inv_d[0];
                    *mat_at(&A, u, u - 1) = -inv_d[0];
                } else if (j == 0) {
                    *mat_at(&A, u, u) = inv_d[1];
                    *mat_at(&A, u, u + ni) = -inv_d[1];
                } else if (j == nj - 1) {
                    *mat_at(&A, u, u) = inv_d[1];
                    *mat_at(&A, u, u - ni) = -inv_d[1];
                } else if (k == 0) {
                    *mat_at(&A, u, u) = inv_d[2];
                    *mat_at(&A, u, u + ni * nj) = -inv_d[2];
                } else if (k == nk - 1) {
                    *mat_at(&A, u, u) = inv_d[2];
                    *mat_at(&A, u, u - ni * nj) = -inv_d[2];
                } else {
                    nt[u] = REG;
                    *mat_at(&A, u, u - ni * nj) = cz;
                    *mat_at(&A, u, u - ni) = s.d2y;
                    *mat_at(&A, u, u - 1) = cx;
                    *mat_at(&A, u, u) = -s.twos;
                    *mat_at(&A, u, u + 1) = cx;
                    *mat_at(&A, u, u + ni) = s.d2y;
                    *mat_at(&A, u, u + ni * nj) = cz;
                }
            }
    double* M = (double*)malloc(n * sizeof(double));
    for (int u = 0; u < n; u++) M[u] = 1 / *mat_at(&A, u, u);

    double* phi_ = (double*)malloc(n * sizeof(double));
    double* b = (double*)malloc(n * sizeof(double));
    double* y = (double*)calloc(n, sizeof(double));
    double* P = (double*)calloc(n, sizeof(double));
    double* F = (double*)malloc(n * sizeof(double));
    memcpy(phi_, phi, n * sizeof(double));
    for (int u = 0; u < n; u++) {
        if (nt[u] == NEU)
            b[u] = 0;
        else if (nt[u] == DIR)
            b[u] = phi_[u];
        else
            b[u] = -rho[u] * s.ieps;
    }
    mat7 J;
    mat_alloc(&J, n);
    int converged = 0, it, total = 0, ret = 0;
    double L2 = 0;
    for (it = 0; it < 20; it++) {
        mat_vec(&A, phi_, F);
        for (int u = 0; u < n; u++) F[u] = F[u] - b[u];
        for (int u = 0; u < n; u++)
            if (nt[u] == REG) F[u] -= ORC_QE * n0 * exp((phi_[u] - phi0) / Te0) * s.ieps;
        for (int u = 0; u < n; u++)
            if (nt[u] == REG) P[u] = ORC_QE * n0 / (ORC_EPS0 * Te0) * exp((phi_[u] - phi0) / Te0);
        memcpy(J.col, A.col, (size_t)n * 7 * sizeof(int32_t));
        memcpy(J.a, A.a, (size_t)n * 7 * sizeof(double));
        for (int u = 0; u < n; u++) {
            double* jd = mat_at(&J, u, u);
            *jd = *mat_at(&A, u, u) - P[u];
        }
        int pit = 0;
        int r = pcg_linear(&J, M, y, F, pcg_max_it, tol, &pit, &L2);
        total += pit;
        if (r < 0) {
            ret = -1;
            break;
        }
        if (!r) gs_linear(&J, y, F, 20 * pcg_max_it, tol);
        for (int u = 0; u < n; u++)
            if (nt[u] == DIR) y[u] = 0;
        for (int u = 0; u < n; u++) phi_[u] = phi_[u] - y[u];
        if (vnorm(y, n) < 1e-3) {
            converged = 1;
            break;
        }
    }
    if (ret == 0) memcpy(phi, phi_, n * sizeof(double));
    if (newton_iters) *newton_iters = converged ? it + 1 : it;
    if (pcg_iters_total) *pcg_iters_total = total;
    if (last_l2) *last_l2 = L2;
    mat_free(&A);
    mat_free(&J);
    free(nt);
    free(M);
    free(phi_);
    free(b);
    free(y);
    free(P);
    free(F);
    return ret < 0 ? -1 : converged;
}

/* r = b - A phi on the interior of a Dirichlet-walled box (physical axes) */
void orc_residual_dirichlet(const orc_grid* g, const double* phi, const double* rho, double* r2,
                            double* b2) {
    stencil_t s;
    stencil_init(g, &s);
    size_t sj = g->ni, sk = (size_t)g->ni * g->nj;
    double rr = 0, bb = 0;
    for (int k = 1; k < g->nk - 1; k++)
        for (int j = 1; j < g->nj - 1; j++)
            for (int i = 1; i < g->ni - 1; i++) {
                size_t u = IDX(g, i, j, k);
                double Ap = (phi[u - 1] + phi[u + 1]) * s.d2x + (phi[u - sj] + phi[u + sj]) * s.d2y +
                            (phi[u - sk] + phi[u + sk]) * s.d2z - s.twos * phi[u];
                double bu = -rho[u] * s.ieps;
                rr += (bu - Ap) * (bu - Ap);
                bb += bu * bu;
            }
    *r2 = sqrt(rr);
    *b2 = sqrt(bb);
}

/* dense elimination with partial pivoting on the interior unknowns */
int orc_dense_poisson(const orc_grid* g, double* phi, const double* rho) {
    int mi = g->ni - 2, mj = g->nj - 2, mk = g->nk - 2;
    int m = mi * mj * mk;
    if (m <= 0 || m > 1000) return -1;
    stencil_t s;
    stencil_init(g, &s);
    double* A = (double*)calloc((size_t)m * m, sizeof(double));
    double* b = (double*)malloc(m * sizeof(double));
    for (int k = 0; k < mk; k++)
        for (int j = 0; j < mj; j++)
            for (int i = 0; i < mi; i++) {
                int r = (k * mj + j) * mi + i;
                size_t u = IDX(g, i + 1, j + 1, k + 1);
                b[r] = -rho[u] * s.ieps;
                A[(size_t)r * m + r] = -s.twos;
                int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0},
                    dk[6] = {0, 0, 0, 0, -1, 1};
                double cf[6] = {s.d2x, s.d2x, s.d2y, s.d2y, s.d2z, s.d2z};
                for (int e = 0; e < 6; e++) {
                    int a = i + di[e], bj = j + dj[e], c = k + dk[e];
                    if (a < 0 || a >= mi || bj < 0 || bj >= mj || c < 0 || c >= mk) {
                        b[r] -= cf[e] * phi[IDX(g, a + 1, bj + 1, c + 1)];
                    } else {
                        A[(size_t)r * m + (c * mj + bj) * mi + a] = cf[e];
                    }
                }
            }
    for (int c = 0; c < m; c++) {
        int piv = c;
        for (int r = c + 1; r < m; r++)
            if (fabs(A[(size_t)r * m + c]) > fabs(A[(size_t)piv * m + c])) piv = r;
        if (piv != c) {
            for (int e = 0; e < m; e++) {
                double t = A[(size_t)c * m + e];
                A[(size_t)c * m + e] = A[(size_t)piv * m + e];
                A[(size_t)piv * m + e] = t;
            }
            double t = b[c];
            b[c] = b[piv];
            b[piv] = t;
        }
        for (int r = c + 1; r < m; r++) {
            double f = A[(size_t)r * m + c] / A[(size_t)c * m + c];
            if (f == 0) continue;
            for (int e = c; e < m; e++) A[(size_t)r * m + e] -= f * A[(size_t)c * m + e];
            b[r] -= f * b[c];
        }
    }
    for (int r = m - 1; r >= 0; r--) {
        double acc = b[r];
        for (int e = r + 1; e < m; e++) acc -= A[(size_t)r * m + e] * b[e];
        b[r] = acc / A[(size_t)r * m + r];
    }
    for (int k = 0; k < mk; k++)
        for (int j = 0; j < mj; j++)
            for (int i = 0; i < mi; i++) phi[IDX(g, i + 1, j + 1, k + 1)] = b[(k * mj + j) * mi + i];
    free(A);
    free(b);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* random numbers                                                            */
/* ------------------------------------------------------------------------ */

/* MT19937 (Matsumoto & Nishimura 1998), the engine behind v3/Rnd.cpp:4-7 */
void orc_mt_seed(orc_mt19937* s, uint32_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < 624; i++)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
    s->mti = 624;
}
uint32_t orc_mt_next(orc_mt19937* s) {
    if (s->mti >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = s->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            s->mt[i] = v;
        }
        s->mti = 0;
    }
    uint32_t y = s->mt[s->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}
/* libstdc++ generate_canonical<double,53> over a 32-bit engine: two draws,
 * low word first, divided by 2^64, clamped below 1. */
double orc_mt_uniform(orc_mt19937* s) {
    double sum = (double)orc_mt_next(s);
    sum += (double)orc_mt_next(s) * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}
double orc_u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

void orc_rng_mt(orc_rng* r, uint32_t seed) {
    memset(r, 0, sizeof(*r));
    r->kind = 0;
    orc_mt_seed(&r->mt, seed);
}
void orc_rng_philox(orc_rng* r, uint32_t seed, uint32_t seq, uint32_t tag) {
    memset(r, 0, sizeof(*r));
    r->kind = 1;
    r->key[0] = seed;
    r->key[1] = seq;
    r->tag = tag;
}
/* philox slot s -> block s/2, half s%2 of the 4-word output */
double orc_rng_draw(orc_rng* r, int slot) {
    if (r->kind == 0) return orc_mt_uniform(&r->mt);
    uint32_t ctr[4] = {r->a, r->b, (uint32_t)(slot >> 1), r->tag};
    uint32_t out[4];
    orc_philox4x32(ctr, r->key, out);
    int h = slot & 1;
    return orc_u53(out[2 * h], out[2 * h + 1]);
}

/* ------------------------------------------------------------------------ */
/* loader                                                                    */
/* ------------------------------------------------------------------------ */
#define TWO_PI_TRUE 6.283185307179586

/* v3/Species.cpp:564-599 + :855-869 (mode 0) and :420-434 (addParticle).
 * Draw slots per particle: 0-2 position, 3-13 velocity. */
size_t orc_load_box_thermal(const orc_grid* g, const orc_rect* rects, int nrect, const double* ef,
                            double charge, double mass, double rewind_dt, const double* xc,
                            const double* sides, double num_den, double T, double mpw0,
                            orc_rng* rng, int mode, size_t max_out, double* parts) {
    double box_vol = sides[0] * sides[1] * sides[2];
    double num_micro = num_den * box_vol;
    size_t num_macro = (size_t)(num_micro / mpw0);
    double xmin[3], xmax[3];
    for (int d = 0; d < 3; d++) {
        xmin[d] = xc[d] - sides[d] * 0.5;
        xmax[d] = xc[d] + sides[d] * 0.5;
    }
    size_t out = 0;
    for (size_t i = 0; i < num_macro; i++) {
        rng->a = (uint32_t)i;
        rng->b = (uint32_t)(i >> 32);
        double pos[3], vel[3];
        for (int d = 0; d < 3; d++) pos[d] = xmin[d] + orc_rng_draw(rng, d) * (xmax[d] - xmin[d]);
        if (mode == 0) {
            double vth = sqrt(2 * ORC_KB * T / mass);
            double v[3];
            for (int c = 0; c < 3; c++) {
                double u1 = orc_rng_draw(rng, 3 + 3 * c);
                double u2 = orc_rng_draw(rng, 4 + 3 * c);
                double u3 = orc_rng_draw(rng, 5 + 3 * c);
                v[c] = vth * (u1 + u2 + u3 - 1.5);
            }
            double sp = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            double theta = 2 * ORC_PI * orc_rng_draw(rng, 12);
            double r = -1.0 + 2 * orc_rng_draw(rng, 13);
            double a = sqrt(1 - r * r);
            vel[0] = sp * r;
            vel[1] = sp * (cos(theta) * a);
            vel[2] = sp * (sin(theta) * a);
        } else {
            double sig = sqrt(ORC_KB * T / mass);
            double u1 = orc_rng_draw(rng, 3), u2 = orc_rng_draw(rng, 4);
            double u3 = orc_rng_draw(rng, 5), u4 = orc_rng_draw(rng, 6);
            double r1 = sqrt(-2.0 * log(1.0 - u1)), r2 = sqrt(-2.0 * log(1.0 - u3));
            vel[0] = sig * r1 * cos(TWO_PI_TRUE * u2);
            vel[1] = sig * r1 * sin(TWO_PI_TRUE * u2);
            vel[2] = sig * r2 * cos(TWO_PI_TRUE * u4);
        }
        /* addParticle: v3/Species.cpp:420-434 */
        if (isnan(pos[0]) || isnan(pos[1]) || isnan(pos[2])) continue;
        if (!orc_in_bounds(g, pos)) continue;
        if (orc_in_object(rects, nrect, pos)) continue;
        if (ef) {
            double lc[3], e[3];
            xtol(g, pos, lc);
            orc_gather3(g, ef, lc, e);
            for (int d = 0; d < 3; d++) vel[d] -= e[d] * (charge / mass) * (0.5 * rewind_dt);
        }
        if (out >= max_out) break;
        double* q = parts + 7 * out;
        for (int d = 0; d < 3; d++) {
            q[d] = pos[d];
            q[3 + d] = vel[d];
        }
        q[6] = mpw0;
        out++;
    }
    return out;
}

/* ------------------------------------------------------------------------ */
/* MC_MEX_Ionization                                                         */
/* ------------------------------------------------------------------------ */

/* v3/Interactions.cpp:541-558 (std::map lower_bound + linear interpolation) */
double orc_sigma_el(const orc_mcc_params* p, double E) {
    int lo = 0, hi = p->nsig; /* first index with key >= E */
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (p->sig_E[mid] < E)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo == 0) return p->sig_v[0];
    if (lo == p->nsig) return p->sig_v[p->nsig - 1];
    double x1 = p->sig_E[lo - 1], x2 = p->sig_E[lo], y1 = p->sig_v[lo - 1], y2 = p->sig_v[lo];
    return y1 + (E - x1) * (y2 - y1) / (x2 - x1);
}

/* v3/Interactions.cpp:559-566 */
double orc_sigma_iz(const orc_mcc_params* p, double E) {
    if (E <= p->E_ion_J / ORC_QE) return 0;
    return p->c0 * log(E / p->c1) / E * exp(-p->c2 / E);
}

static void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* v3/Interactions.cpp:845-870 (non-normalised basis kept as in the source) */
static void new_velocity_electron(double* vel, double E, const double* u, double U1, double U2) {
    double cos_ksi = (2 + E - 2 * pow(1 + E, U1)) / E;
    double sin_ksi = sqrt(1 - cos_ksi * cos_ksi);
    double phi = 2 * ORC_PI * U2;
    double v_mag = sqrt(E * (2 * ORC_QE / ORC_ME));
    static const double ix[3] = {1, 0, 0};
    double ic[3], uic[3];
    cross3(ix, u, ic);
    cross3(u, ic, uic);
    double sp = sin(phi), cp = cos(phi);
    for (int d = 0; d < 3; d++) {
        double v = u[d] * cos_ksi + ic[d] * sin_ksi * sp + uic[d] * sin_ksi * cp;
        vel[d] = v * v_mag;
    }
}

/* v3/Interactions.cpp:884-947 (IONIZE_1 variant).  Draw slots 3..8. */
int orc_mcc_collide(const orc_mcc_params* p, orc_rng* rng, double* vel_neu, double* vel_ele,
                    double* vel_new, double sigma_coll) {
    double m_red = p->m_neutral * p->m_electron / (p->m_neutral + p->m_electron);
    double sum_mass = p->m_neutral + p->m_electron;
    double E_ion_eV = p->E_ion_J / ORC_QE;
    double vr[3] = {vel_neu[0] - vel_ele[0], vel_neu[1] - vel_ele[1], vel_neu[2] - vel_ele[2]};
    double vrm = sqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
    double E_rel_J = 0.5 * m_red * vrm * vrm;
    double P = orc_sigma_iz(p, E_rel_J / ORC_QE) / sigma_coll;
    if (orc_rng_draw(rng, 3) <= P) {
        double vem = sqrt(vel_ele[0] * vel_ele[0] + vel_ele[1] * vel_ele[1] + vel_ele[2] * vel_ele[2]);
        double E_inc = vem * vem * (ORC_ME * 0.5 / ORC_QE);
        if (E_inc < E_ion_eV) return 0;
        double E_ej = 10.0 * tan(orc_rng_draw(rng, 4) * atan((E_inc - E_ion_eV) / (2 * p->B_inc)));
        double E_sc = E_inc - E_ion_eV - E_ej;
        if (E_sc < 0) E_sc = 0.000001;
        double inv = 1.0 / vem;
        double u[3] = {vel_ele[0] * inv, vel_ele[1] * inv, vel_ele[2] * inv};
        double U5 = orc_rng_draw(rng, 5), U6 = orc_rng_draw(rng, 6);
        new_velocity_electron(vel_ele, E_sc, u, U5, U6);
        double U7 = orc_rng_draw(rng, 7), U8 = orc_rng_draw(rng, 8);
        new_velocity_electron(vel_new, E_ej, u, U7, U8);
        return 1;
    }
    double inv = 1.0 / sum_mass;
    double cm[3];
    for (int d = 0; d < 3; d++)
        cm[d] = (vel_neu[d] * p->m_neutral + vel_ele[d] * p->m_electron) * inv;
    double cos_ksi = 2 * orc_rng_draw(rng, 4) - 1;
    double sin_ksi = sqrt(1 - cos_ksi * cos_ksi);
    double eps = 2 * ORC_PI * orc_rng_draw(rng, 5);
    double v2[3] = {vrm * cos_ksi, vrm * sin_ksi * cos(eps), vrm * sin_ksi * sin(eps)};
    double f = p->m_neutral / sum_mass;
    for (int d = 0; d < 3; d++) vel_ele[d] = cm[d] - v2[d] * f;
    return 0;
}

static int push_part(double* arr, size_t* n, size_t cap, const double* pos, const double* vel,
                     double w) {
    if (*n >= cap) return -1;
    double* q = arr + 7 * (*n);
    for (int d = 0; d < 3; d++) {
        q[d] = pos[d];
        q[3 + d] = vel[d];
    }
    q[6] = w;
    (*n)++;
    return 0;
}

/* v3/Interactions.cpp:600-762.  Defined semantics for the electron-heavier
 * branch: the differential electron carries w_e - w_n and an ionised neutral
 * is removed at the end of the call (the source's swap-with-back() is
 * undefined once several neutrals are removed).  Draw slots: 0 neutral
 * pick, 1 electron pick, 2 accept, 3.. collide. */
int orc_mcc_apply(const orc_grid* g, const orc_mcc_params* p, double* wsv_max, orc_rng* rng,
                  double* neu, size_t* nn, size_t cap_n, double* ion, size_t* nion, size_t cap_i,
                  double* ele, size_t* ne, size_t cap_e, orc_mcc_stats* st) {
    int nc = (g->ni - 1) * (g->nj - 1) * (g->nk - 1);
    int32_t* offn = (int32_t*)malloc((size_t)(nc + 1) * sizeof(int32_t));
    int32_t* offe = (int32_t*)malloc((size_t)(nc + 1) * sizeof(int32_t));
    int32_t* idxn = (int32_t*)malloc((*nn + 1) * sizeof(int32_t));
    int32_t* idxe = (int32_t*)malloc((*ne + 1) * sizeof(int32_t));
    orc_sort_indexes(g, neu, *nn, offn, idxn);
    orc_sort_indexes(g, ele, *ne, offe, idxe);
    double E_pre = 0.5 * (p->m_neutral * p->m_electron / (p->m_neutral + p->m_electron)) / ORC_QE;
    double inv_dv = 1 / p->dV;
    double wsv_tmp = 0;
    long long ncoll = 0, niz = 0, ngroups = 0;
    int err = 0;
    unsigned char* dead = (unsigned char*)calloc(cap_n + 1, 1);
    for (int c = 0; c < nc && !err; c++) {
        int Ne = offe[c + 1] - offe[c];
        if (Ne <= 0) continue;
        int Nn = offn[c + 1] - offn[c];
        if (Nn <= 0) continue;
        int nfrac_n = Nn;
        double frac = (double)(Nn * Ne) * (*wsv_max) * p->dt * inv_dv;
        int n_groups = (int)(frac + 0.5);
        if (n_groups > Nn) n_groups = Nn - 1;
        (void)nfrac_n;
        /* per-cell pools (copies, may grow) */
        int capn = Nn + (n_groups > 0 ? n_groups : 0) + 1;
        int cape = Ne + (n_groups > 0 ? n_groups : 0) + 1;
        int32_t* pn = (int32_t*)malloc(capn * sizeof(int32_t));
        int32_t* pe = (int32_t*)malloc(cape * sizeof(int32_t));
        memcpy(pn, idxn + offn[c], Nn * sizeof(int32_t));
        memcpy(pe, idxe + offe[c], Ne * sizeof(int32_t));
        int pe_len = Ne;
        for (int gi = 0; gi < n_groups; gi++) {
            rng->a = (uint32_t)c;
            rng->b = (uint32_t)gi;
            ngroups++;
            int ns = (int)(0.0 + orc_rng_draw(rng, 0) * ((double)Nn - 0.0));
            int p_neu = pn[ns];
            int es = (int)(0.0 + orc_rng_draw(rng, 1) * ((double)Ne - 0.0));
            int p_ele = pe[es];
            double* qn = neu + 7 * (size_t)p_neu;
            double* qe = ele + 7 * (size_t)p_ele;
            double vr[3] = {qn[3] - qe[3], qn[4] - qe[4], qn[5] - qe[5]};
            double v_rel = sqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
            double E_rel = E_pre * v_rel * v_rel;
            double sigma = orc_sigma_el(p, E_rel);
            double Wg = qn[6], Wl = qe[6];
            if (Wg < Wl) {
                double t = Wg;
                Wg = Wl;
                Wl = t;
            }
            double wsv = Wg * sigma * v_rel;
            if (wsv > wsv_tmp) wsv_tmp = wsv;
            double P = wsv / (*wsv_max);
            if (orc_rng_draw(rng, 2) < P) {
                ncoll++;
                if (qn[6] > qe[6]) {
                    qn[6] -= qe[6];
                    double vsplit[3] = {qn[3], qn[4], qn[5]};
                    double vnew[3] = {0, 0, 0};
                    int iz = orc_mcc_collide(p, rng, vsplit, qe + 3, vnew, sigma);
                    if (iz) {
                        niz++;
                        double posn[3] = {qn[0], qn[1], qn[2]};
                        if (push_part(ion, nion, cap_i, posn, vsplit, Wl) ||
                            push_part(ele, ne, cap_e, posn, vnew, Wl)) {
                            err = 1;
                            break;
                        }
                    } else {
                        double posn[3] = {qn[0], qn[1], qn[2]}, veln[3] = {qn[3], qn[4], qn[5]};
                        if (push_part(neu, nn, cap_n, posn, veln, qe[6])) {
                            err = 1;
                            break;
                        }
                        qn = neu + 7 * (size_t)p_neu;
                        if (p->pool_append) {
                            pn[Nn] = (int32_t)(*nn - 1);
                            Nn++;
                        }
                    }
                } else if (qn[6] < qe[6]) {
                    double w_e_old = qe[6];
                    qe[6] = qn[6];
                    double pose[3] = {qe[0], qe[1], qe[2]}, vele[3] = {qe[3], qe[4], qe[5]};
                    if (push_part(ele, ne, cap_e, pose, vele, w_e_old - qn[6])) {
                        err = 1;
                        break;
                    }
                    qe = ele + 7 * (size_t)p_ele;
                    if (pe_len < cape) pe[pe_len++] = (int32_t)(*ne - 1);
                    double vnew[3] = {0, 0, 0};
                    int iz = orc_mcc_collide(p, rng, qn + 3, qe + 3, vnew, sigma);
                    if (iz) {
                        niz++;
                        double posn[3] = {qn[0], qn[1], qn[2]}, veln[3] = {qn[3], qn[4], qn[5]};
                        if (push_part(ion, nion, cap_i, posn, veln, Wl) ||
                            push_part(ele, ne, cap_e, posn, vnew, Wl)) {
                            err = 1;
                            break;
                        }
                        Nn--;
                        pn[ns] = pn[Nn];
                        dead[p_neu] = 1;
                    }
                }
            }
        }
        free(pn);
        free(pe);
    }
    if (ncoll) {
        *wsv_max = wsv_tmp;
        size_t out = 0;
        for (size_t q = 0; q < *nn; q++) {
            if (dead[q]) continue;
            if (out != q) memmove(neu + 7 * out, neu + 7 * q, 7 * sizeof(double));
            out++;
        }
        *nn = out;
    }
    if (st) {
        st->collisions = ncoll;
        st->ionizations = niz;
        st->groups = ngroups;
        st->wsv_max_step = wsv_tmp;
    }
    free(dead);
    free(offn);
    free(offe);
    free(idxn);
    free(idxe);
    return err ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* threaded CPU baseline push (v3/Species.cpp:258-355, v3/Config.cpp:79-99) */
/* ------------------------------------------------------------------------ */
typedef struct {
    const orc_grid* g;
    const double* ef;
    double* parts;
    size_t start, stop;
    double qmdt, dt;
    size_t* del;
    size_t ndel;
} push_task;

static void* push_worker(void* arg) {
    push_task* t = (push_task*)arg;
    t->ndel = 0;
    for (size_t p = t->start; p < t->stop; p++)
        if (!push_one(t->g, t->ef, t->parts + 7 * p, t->qmdt, t->dt, NULL, 0))
            t->del[t->ndel++] = p;
    return NULL;
}
static int cmp_desc(const void* a, const void* b) {
    size_t x = *(const size_t*)a, y = *(const size_t*)b;
    return x < y ? 1 : (x > y ? -1 : 0);
}
/* v3/Species.cpp:263-310.  The per-thread deletion buffers reach the
 * workers through ThreadPool::AddTask's std::bind (ThreadPool.h:63, :77-78),
 * which stores copies: the workers fill the copies, the caller's buffers
 * stay empty, and the reference removes nothing (remove == 0 restates
 * that).  remove != 0 applies the removal the code intends: the buffers
 * merged, sorted descending, each slot swapped with the last. */
size_t orc_push_threaded(const orc_grid* g, const double* ef, double* parts, size_t n,
                         double charge, double mass, double dt, int nthreads, int remove) {
    if (nthreads < 1) nthreads = 1;
    push_task* tasks = (push_task*)calloc(nthreads, sizeof(push_task));
    pthread_t* th = (pthread_t*)calloc(nthreads, sizeof(pthread_t));
    size_t base = n / nthreads, more = n % nthreads, start = 0;
    for (int t = 0; t < nthreads; t++) {
        size_t len = base + ((size_t)t < more ? 1 : 0);
        tasks[t] = (push_task){g, ef, parts, start, start + len, dt * charge / mass, dt, NULL, 0};
        tasks[t].del = (size_t*)malloc((len + 1) * sizeof(size_t));
        start += len;
        pthread_create(&th[t], NULL, push_worker, &tasks[t]);
    }
    size_t ndel = 0;
    for (int t = 0; t < nthreads; t++) {
        pthread_join(th[t], NULL);
        if (remove) ndel += tasks[t].ndel;
    }
    if (ndel) {
        size_t* all = (size_t*)malloc(ndel * sizeof(size_t));
        size_t o = 0;
        for (int t = 0; t < nthreads; t++)
            for (size_t e = 0; e < tasks[t].ndel; e++) all[o++] = tasks[t].del[e];
        qsort(all, ndel, sizeof(size_t), cmp_desc);
        size_t last = n - 1;
        for (size_t e = 0; e < ndel; e++) {
            memcpy(parts + 7 * all[e], parts + 7 * last, 7 * sizeof(double));
            last--;
        }
        free(all);
    }
    for (int t = 0; t < nthreads; t++) free(tasks[t].del);
    free(tasks);
    free(th);
    return n - ndel;
}

/* ------------------------------------------------------------------------ */
/* object surfaces: v3/Object.cpp:271-317, v3/Species.cpp:170-256, :835-861  */
/* ------------------------------------------------------------------------ */

/* v3/Object.cpp:310-316 (Rectangle::find_n): n += outward normal of side s */
static void rect_add_normal(int side, double* n) {
    static const double nrm[6][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1},
                                     {1, 0, 0},  {0, 1, 0},  {0, 0, 1}};
    for (int d = 0; d < 3; d++) n[d] += nrm[side][d];
}

/* v3/Object.cpp:271-308 (Rectangle::lineIntersect; x2 inside the box).
 * Entry time from the slab method, the entry side's outward normal; within
 * 1e-6 of another side's time the normals are summed and normalised. */
void orc_rect_line_intersect(const orc_rect* r, const double* x1, const double* x2, double* t_entry,
                             double* pos, double* n) {
    double A[3], tmin[3], tmax[3];
    int side[3] = {0, 1, 2};
    for (int d = 0; d < 3; d++) {
        double xmin = r->c[d] - r->half[d], xmax = r->c[d] + r->half[d]; /* Object.cpp:205-206 */
        A[d] = x2[d] - x1[d];
        tmin[d] = (xmin - x1[d]) / A[d];
        tmax[d] = (xmax - x1[d]) / A[d];
    }
    for (int d = 0; d < 3; d++)
        if (tmin[d] > tmax[d]) {
            double t = tmin[d];
            tmin[d] = tmax[d];
            tmax[d] = t;
            side[d] += 3;
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
    n[0] = n[1] = n[2] = 0;
    rect_add_normal(side[j], n);
    if (k != 0) {
        if (k < 2) {
            rect_add_normal(side[t], n);
        } else {
            rect_add_normal(side[2], n);
            rect_add_normal(side[1], n);
        }
        double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]); /* Vec3.h:102-108 */
        double inv = 1.0 / len;
        for (int d = 0; d < 3; d++) n[d] *= inv;
    }
    *t_entry = te;
    for (int d = 0; d < 3; d++) pos[d] = x1[d] + te * (x2[d] - x1[d]);
}

/* v3/Species.cpp:855-861 (sampleVth) then :835-853 (sampleReflectedVelocity):
 * 300 K wall, full accommodation, the diffuse direction built from
 * t1 = n x e_x (n x e_y when n . e_x == 0) and t2 = n x t1 as in the source
 * (not normalised; sin(psi) t2 carries no sin(theta)).  Draw slots s0 ..
 * s0 + 10: nine for the thermal speed, then sin(theta), then psi. */
static void reflect_velocity(double mass, double vmag1, const double* n, orc_rng* rng, int s0,
                             double* out) {
    double vth0 = sqrt(2 * ORC_KB * 300.0 / mass);
    double v[3];
    for (int c = 0; c < 3; c++) {
        double u1 = orc_rng_draw(rng, s0 + 3 * c);
        double u2 = orc_rng_draw(rng, s0 + 3 * c + 1);
        double u3 = orc_rng_draw(rng, s0 + 3 * c + 2);
        v[c] = vth0 * (u1 + u2 + u3 - 1.5);
    }
    double v_th = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double vmag2 = vmag1 + 1.0 * (v_th - vmag1);
    double sin_theta = orc_rng_draw(rng, s0 + 9);
    double cos_theta = sqrt(1 - sin_theta * sin_theta);
    double psi = 2 * ORC_PI * orc_rng_draw(rng, s0 + 10);
    static const double ex[3] = {1, 0, 0}, ey[3] = {0, 1, 0};
    double t1[3], t2[3];
    if (n[0] * 1.0 + n[1] * 0.0 + n[2] * 0.0 != 0)
        cross3(n, ex, t1);
    else
        cross3(n, ey, t1);
    cross3(n, t1, t2);
    double a = sin_theta * cos(psi), b = sin(psi);
    for (int d = 0; d < 3; d++) out[d] = ((t1[d] * a + t2[d] * b) + n[d] * cos_theta) * vmag2;
}

/* v3/Species.cpp:170-256 (advanceNoSputteringSerial) for a species of the
 * given charge: kick, then the drift loop -- leaving the box removes; ending
 * in an object (the first containing the point) moves the particle to the
 * entry point; a neutral (charge 0) reflects diffusely and drifts on for the
 * remaining fraction of the step, an ion is removed and emits
 * int(w / mpw0_n + U) neutrals there (addParticle rejects a point that still
 * tests inside the object unless accept_on_surface); more than 20 bounces
 * remove.  Draws per (particle, bounce): slot 0 the emission count, then 11
 * per sampled velocity (reflect_velocity); in Philox mode the counter words
 * are (input index, bounce).  stable != 0 keeps the survivors' order (and
 * the input index equals the position); otherwise the reference's
 * swap-with-last removal.  stats (may be NULL): {object hits, neutrals
 * emitted, emissions rejected, particles removed}.  Returns survivors, or
 * (size_t)-1 when the neutral capacity is exceeded. */
size_t orc_push_nonelectron(const orc_grid* g, const double* ef, double* parts, size_t n,
                            double charge, double mass, double dt, const orc_rect* rects, int nrect,
                            double* neu, size_t* n_neu, size_t cap_neu, double mpw0_n,
                            orc_rng* rng, int accept_on_surface, int stable, long long* stats) {
    double qmdt = dt * charge / mass;
    long long st[4] = {0, 0, 0, 0};
    size_t np = n, out = 0;
    for (size_t p = 0; p < np; p++) {
        double* q = parts + 7 * p;
        double lc[3], e[3];
        xtol(g, q, lc);
        orc_gather3(g, ef, lc, e);
        for (int d = 0; d < 3; d++) q[3 + d] += e[d] * qmdt;
        double t_rem = 1;
        int nb = 0, alive = 1;
        while (t_rem > 0) {
            if (++nb > 20) {
                alive = 0;
                break;
            }
            rng->a = (uint32_t)p;
            rng->b = (uint32_t)nb;
            double old[3] = {q[0], q[1], q[2]};
            for (int d = 0; d < 3; d++) q[d] += q[3 + d] * t_rem * dt;
            int obj = orc_in_object(rects, nrect, q);
            if (!orc_in_bounds(g, q)) {
                alive = 0;
                break;
            }
            if (obj) {
                double tp, nrm[3];
                st[0]++;
                orc_rect_line_intersect(&rects[obj - 1], old, q, &tp, q, nrm);
                double vm = sqrt(q[3] * q[3] + q[4] * q[4] + q[5] * q[5]);
                if (charge == 0) {
                    reflect_velocity(mass, vm, nrm, rng, 0, q + 3);
                } else {
                    int mp = (int)(q[6] / mpw0_n + orc_rng_draw(rng, 0));
                    for (int m = 0; m < mp; m++) {
                        double v[3];
                        reflect_velocity(mass, vm, nrm, rng, 1 + 11 * m, v);
                        /* neutrals.addParticle (Species.cpp:420-434; no rewind at charge 0) */
                        if (isnan(q[0]) || isnan(q[1]) || isnan(q[2]) || isnan(v[0]) ||
                            isnan(v[1]) || isnan(v[2]) || !orc_in_bounds(g, q) ||
                            (!accept_on_surface && orc_in_object(rects, nrect, q))) {
                            st[2]++;
                            continue;
                        }
                        if (*n_neu >= cap_neu) return (size_t)-1;
                        double* r = neu + 7 * (*n_neu);
                        for (int d = 0; d < 3; d++) {
                            r[d] = q[d];
                            r[3 + d] = v[d];
                        }
                        r[6] = mpw0_n;
                        (*n_neu)++;
                        st[1]++;
                    }
                    alive = 0;
                    break;
                }
                t_rem *= (1 - tp);
                continue;
            }
            t_rem = 0;
        }
        if (!alive) st[3]++;
        if (stable) {
            if (alive) {
                if (out != p) memmove(parts + 7 * out, q, 7 * sizeof(double));
                out++;
            }
        } else if (!alive) {
            memcpy(q, parts + 7 * (np - 1), 7 * sizeof(double));
            np--;
            p--;
        }
    }
    if (stats)
        for (int s = 0; s < 4; s++) stats[s] = st[s];
    return stable ? out : np;
}

/* ------------------------------------------------------------------------ */
/* variable-weight merge: v3/Species.cpp:987-1145                            */
/* ------------------------------------------------------------------------ */
#define MERGE_BINS 15 /* m_vel_grid_n, Species.h:129 */

typedef struct {
    int bin; /* (i * 15 + j) * 15 + k */
    int pos; /* place in the cell's list */
} bin_item;
static int cmp_bin_item(const void* a, const void* b) {
    const bin_item *x = (const bin_item*)a, *y = (const bin_item*)b;
    if (x->bin != y->bin) return x->bin < y->bin ? -1 : 1;
    return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

/* v3/Species.cpp:1017-1145.  gpu_mode 0 restates the reference: velocity
 * range over the cell and particles[0] (:1020-1033), bins of width
 * range / 15 after widening the maximum by 0.1% (:1043-1067), per bin of
 * more than two particles the pair v = mean +- sqrt(variance) per axis with
 * half the weight each (:1078-1103), positions of two distinct random
 * members (:1127-1133), appended through addParticle (half-step rewind,
 * :1141-1142), a negative variance below -1e-4 stops the clamping loop
 * (:1091-1098), zero-weight particles swap-removed at the end (:1149-1156);
 * mt19937 draws.  gpu_mode 1 is the device's defined variant: the range
 * over the cell only, every negative variance clamped to 0, the pair
 * written in place of the bin's first two members (in cell order) and the
 * rest removed, Philox draws with counter words (cell, bin), survivors kept
 * in order.  A bin index past 14 (a component equal to a zero maximum) or a
 * NaN one (an all-zero component) is clamped into [0, 14] in both modes
 * (undefined behaviour in the source).  parts must have room for cap
 * particles.  stats (may be NULL): {bins merged, particles in them}.
 * Returns 0, or -1 when the capacity is exceeded. */
int orc_merge(const orc_grid* g, const double* ef, double* parts, size_t* n, size_t cap,
              double charge, double mass, double dt, orc_rng* rng, int gpu_mode, long long* stats) {
    size_t n0 = *n;
    int nc = (g->ni - 1) * (g->nj - 1) * (g->nk - 1);
    int32_t* off = (int32_t*)malloc((size_t)(nc + 1) * sizeof(int32_t));
    int32_t* idx = (int32_t*)malloc((n0 + 1) * sizeof(int32_t));
    orc_sort_indexes(g, parts, n0, off, idx);
    long long st[2] = {0, 0};
    int err = 0;
    bin_item* items = (bin_item*)malloc((n0 + 1) * sizeof(bin_item));
    for (int c = 0; c < nc && !err; c++) {
        int cn = off[c + 1] - off[c];
        if (cn < 10) continue;
        const int32_t* cell = idx + off[c];
        double vmin[3], vmax[3];
        for (int d = 0; d < 3; d++)
            vmin[d] = vmax[d] = gpu_mode ? parts[7 * (size_t)cell[0] + 3 + d] : parts[3 + d];
        for (int q = 0; q < cn; q++) {
            const double* v = parts + 7 * (size_t)cell[q] + 3;
            for (int d = 0; d < 3; d++) {
                if (v[d] < vmin[d])
                    vmin[d] = v[d];
                else if (v[d] > vmax[d])
                    vmax[d] = v[d];
            }
        }
        double dv[3];
        for (int d = 0; d < 3; d++) {
            if (vmax[d] > 0)
                vmax[d] *= 1.001;
            else
                vmax[d] *= 0.999;
            dv[d] = (vmax[d] - vmin[d]) / (double)MERGE_BINS;
        }
        for (int q = 0; q < cn; q++) {
            const double* v = parts + 7 * (size_t)cell[q] + 3;
            int b[3];
            for (int d = 0; d < 3; d++) {
                double x = (v[d] - vmin[d]) / dv[d];
                int bi = x == x ? (x < (double)MERGE_BINS ? (int)x : MERGE_BINS - 1) : 0;
                b[d] = bi < 0 ? 0 : bi;
            }
            items[q].bin = (b[0] * MERGE_BINS + b[1]) * MERGE_BINS + b[2];
            items[q].pos = q;
        }
        qsort(items, cn, sizeof(bin_item), cmp_bin_item);
        for (int s = 0; s < cn && !err;) {
            int e = s;
            while (e < cn && items[e].bin == items[s].bin) e++;
            int len = e - s;
            if (len > 2) {
                st[0]++;
                st[1] += len;
                double tw = 0, tm[3] = {0, 0, 0}, te[3] = {0, 0, 0};
                for (int q = s; q < e; q++) {
                    double* pp = parts + 7 * (size_t)cell[items[q].pos];
                    double w = pp[6];
                    tw += w;
                    for (int d = 0; d < 3; d++) tm[d] += pp[3 + d] * w;
                    for (int d = 0; d < 3; d++) te[d] += (pp[3 + d] * pp[3 + d]) * w;
                    pp[6] = 0;
                }
                double wa = tw * 0.5, wb = tw * 0.5;
                double inv = 1.0 / tw;
                for (int d = 0; d < 3; d++) tm[d] *= inv;
                for (int d = 0; d < 3; d++) te[d] *= inv;
                for (int d = 0; d < 3; d++) te[d] -= tm[d] * tm[d];
                for (int d = 0; d < 3; d++) {
                    if (te[d] < 0) {
                        if (te[d] < -10e-5 && !gpu_mode) break;
                        te[d] = 0;
                    }
                }
                double va[3], vb[3];
                for (int d = 0; d < 3; d++) {
                    double sq = sqrt(te[d]);
                    va[d] = tm[d] + sq;
                    vb[d] = tm[d] - sq;
                }
                rng->a = (uint32_t)c;
                rng->b = (uint32_t)items[s].bin;
                int slot = 0;
                int ia = (int)(0.0 + orc_rng_draw(rng, slot++) * ((double)len - 0.0));
                int ib = (int)(0.0 + orc_rng_draw(rng, slot++) * ((double)len - 0.0));
                while (ib == ia) ib = (int)(0.0 + orc_rng_draw(rng, slot++) * ((double)len - 0.0));
                double pa[3], pb[3];
                const double* qa = parts + 7 * (size_t)cell[items[s + ia].pos];
                const double* qb = parts + 7 * (size_t)cell[items[s + ib].pos];
                for (int d = 0; d < 3; d++) {
                    pa[d] = qa[d];
                    pb[d] = qb[d];
                }
                double* dst[2] = {NULL, NULL};
                if (gpu_mode) {
                    dst[0] = parts + 7 * (size_t)cell[items[s].pos];
                    dst[1] = parts + 7 * (size_t)cell[items[s + 1].pos];
                }
                const double* pos2[2] = {pa, pb};
                double* vel2[2] = {va, vb};
                double w2[2] = {wa, wb};
                for (int m = 0; m < 2; m++) {
                    double* v = vel2[m];
                    const double* x = pos2[m];
                    /* addParticle (v3/Species.cpp:420-434) */
                    if (isnan(x[0]) || isnan(x[1]) || isnan(x[2]) || isnan(v[0]) || isnan(v[1]) ||
                        isnan(v[2]) || !orc_in_bounds(g, x))
                        continue;
                    if (ef) {
                        double lc[3], E[3];
                        xtol(g, x, lc);
                        orc_gather3(g, ef, lc, E);
                        for (int d = 0; d < 3; d++) v[d] -= E[d] * (charge / mass) * (0.5 * dt);
                    }
                    double* r = dst[m];
                    if (!r) {
                        if (*n >= cap) {
                            err = 1;
                            break;
                        }
                        r = parts + 7 * (*n);
                        (*n)++;
                    }
                    for (int d = 0; d < 3; d++) {
                        r[d] = x[d];
                        r[3 + d] = v[d];
                    }
                    r[6] = w2[m];
                }
            }
            s = e;
        }
    }
    free(items);
    free(off);
    free(idx);
    if (err) return -1;
    size_t np = *n;
    if (gpu_mode) {
        size_t out = 0;
        for (size_t p = 0; p < np; p++) {
            if (parts[7 * p + 6] == 0) continue;
            if (out != p) memmove(parts + 7 * out, parts + 7 * p, 7 * sizeof(double));
            out++;
        }
        np = out;
    } else {
        for (size_t p = 0; p < np; p++) {
            if (parts[7 * p + 6] == 0) {
                memcpy(parts + 7 * p, parts + 7 * (np - 1), 7 * sizeof(double));
                np--;
                p--;
            }
        }
    }
    *n = np;
    if (stats) {
        stats[0] = st[0];
        stats[1] = st[1];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* DSMC_MEX: v3/Interactions.cpp:143-284 (equal-weight DSMC, Bird's VHS)     */
/* ------------------------------------------------------------------------ */
/* v3/Interactions.cpp:153-156, :175-177 */
static double vhs_sigma(double m_reduced, double v_rel) {
    double c0 = 4.07e-10, c1 = 0.77, c2 = 2 * ORC_KB * 273.15 / m_reduced, c3 = tgamma(2.5 - c1);
    return ORC_PI * c0 * c0 * pow(c2 / (v_rel * v_rel), c1 - 0.5) / c3;
}

/* v3/Interactions.cpp:264-284 */
static void dsmc_collide(double m1, double m2, double* v1, double* v2, double U1, double U2) {
    double sum = m1 + m2, inv = 1.0 / sum;
    double cm[3], vr[3];
    for (int d = 0; d < 3; d++) {
        cm[d] = (v1[d] * m1 + v2[d] * m2) * inv;
        vr[d] = v1[d] - v2[d];
    }
    double vm = sqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
    double cos_ksi = 2 * U1 - 1;
    double sin_ksi = sqrt(1 - cos_ksi * cos_ksi);
    double eps = 2 * ORC_PI * U2;
    vr[0] = vm * cos_ksi;
    vr[1] = vm * sin_ksi * cos(eps);
    vr[2] = vm * sin_ksi * sin(eps);
    double f2 = m2 / sum, f1 = m1 / sum;
    for (int d = 0; d < 3; d++) {
        v1[d] = cm[d] + vr[d] * f2;
        v2[d] = cm[d] - vr[d] * f1;
    }
}

/* v3/Interactions.cpp:181-259: applyOneSpecies (p2 == NULL) or
 * applyTwoSpecies.  Per cell: n_groups = (int)(0.5 N1 N2 mpw0 svmax dt/dV +
 * 0.5) (N1 = N2 = N for one species, which needs N >= 2); random pairs
 * (distinct within one species), accept if P = sigma v / svmax > U, then
 * an isotropic scattering in the centre of mass; svmax <- the step maximum
 * when anything collided.  mt19937 order: picks, accept, cos, eps.  Philox
 * slots per (cell, group): 0 accept, 1 cos, 2 eps, 3 first pick, 4.. the
 * second pick's draws.  Returns collisions; *groups the NTC groups. */
long long orc_dsmc_apply(const orc_grid* g, double* p1, size_t n1, double mass1, double* p2,
                         size_t n2, double mass2, double mpw0, double dt, double* svmax,
                         orc_rng* rng, long long* groups) {
    int nc = (g->ni - 1) * (g->nj - 1) * (g->nk - 1);
    int one = p2 == NULL;
    if (one) {
        p2 = p1;
        n2 = n1;
        mass2 = mass1;
    }
    double m_red = mass1 * mass2 / (mass1 + mass2);
    double dv = g->dx[0] * g->dx[1] * g->dx[2];
    int32_t* off1 = (int32_t*)malloc((size_t)(nc + 1) * sizeof(int32_t));
    int32_t* idx1 = (int32_t*)malloc((n1 + 1) * sizeof(int32_t));
    int32_t* off2 = off1;
    int32_t* idx2 = idx1;
    orc_sort_indexes(g, p1, n1, off1, idx1);
    if (!one) {
        off2 = (int32_t*)malloc((size_t)(nc + 1) * sizeof(int32_t));
        idx2 = (int32_t*)malloc((n2 + 1) * sizeof(int32_t));
        orc_sort_indexes(g, p2, n2, off2, idx2);
    }
    double tmp_max = 0;
    long long ncoll = 0, ngr = 0;
    for (int c = 0; c < nc; c++) {
        int N1 = off1[c + 1] - off1[c], N2 = off2[c + 1] - off2[c];
        if (one ? N1 < 2 : (N1 < 1 || N2 < 1)) continue;
        double frac = 0.5 * N1 * N2 * mpw0 * (*svmax) * dt / dv;
        int ng = (int)(frac + 0.5);
        for (int gi = 0; gi < ng; gi++) {
            ngr++;
            rng->a = (uint32_t)c;
            rng->b = (uint32_t)gi;
            int slot = 3;
            int a = (int)(orc_rng_draw(rng, slot++) * N1);
            int b = (int)(orc_rng_draw(rng, slot++) * N2);
            while (one && a == b) b = (int)(orc_rng_draw(rng, slot++) * N2);
            double* va = p1 + 7 * (size_t)idx1[off1[c] + a] + 3;
            double* vb = p2 + 7 * (size_t)idx2[off2[c] + b] + 3;
            double vr[3] = {va[0] - vb[0], va[1] - vb[1], va[2] - vb[2]};
            double v_rel = sqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
            double sv = vhs_sigma(m_red, v_rel) * v_rel;
            if (sv > tmp_max) tmp_max = sv;
            double P = sv / (*svmax);
            if (P > orc_rng_draw(rng, 0)) {
                ncoll++;
                double U1 = orc_rng_draw(rng, 1);
                double U2 = orc_rng_draw(rng, 2);
                dsmc_collide(mass1, mass2, va, vb, U1, U2);
            }
        }
    }
    if (ncoll) *svmax = tmp_max;
    free(off1);
    free(idx1);
    if (!one) {
        free(off2);
        free(idx2);
    }
    if (groups) *groups = ngr;
    return ncoll;
}

/* ------------------------------------------------------------------------ */
/* sources: v3/Source.cpp:3-185 (cold and warm beams through a face)          */
/* ------------------------------------------------------------------------ */
/* Source::Source (:3-28): the injected count per call is
 * (int)(den v_drift A dt / mpw0 + U), A the face area; ColdBeamSource and
 * WarmBeamSource (:32-185): positions uniform on the face (a + face at
 * x0 + L - dx, :47-48), velocity v_drift along the inward normal, plus
 * sampleV3th(T) for a warm beam (T > 0, drawn before the position);
 * addParticle (Species.cpp:420-434) rejects points outside the box or in an
 * object and rewinds v by half a step in E.  face: 0 x-, 1 x+, 2 y-, 3 y+,
 * 4 z-, 5 z+.  Draws, mt19937: the count, then per particle (warm) the
 * eleven of sampleV3th and the two position draws.  Philox: the count at
 * counter (0xFFFFFFFF, 0xFFFFFFFF) slot 0, particle i at (i, 0) slots 0-10
 * (warm) and 11-12 (position).  Returns particles appended, or (size_t)-1
 * past max_out. */
size_t orc_source_sample(const orc_grid* g, const orc_rect* rects, int nrect, const double* ef,
                         double charge, double mass, double mpw0, int face, double v_drift,
                         double den, double T, double dt, orc_rng* rng, double* parts,
                         size_t max_out) {
    double L[3];
    for (int d = 0; d < 3; d++) L[d] = g->dx[d] * ((d == 0 ? g->ni : d == 1 ? g->nj : g->nk) - 1);
    int ax = face / 2, plus = face & 1;
    int t0 = ax == 0 ? 1 : 0, t1 = ax == 2 ? 1 : 2;
    double A = L[t0] * L[t1];
    double num_micro = den * v_drift * A * dt;
    if (plus) L[ax] -= g->dx[ax];
    rng->a = 0xFFFFFFFFu;
    rng->b = 0xFFFFFFFFu;
    int num_macro = (int)(num_micro / mpw0 + orc_rng_draw(rng, 0));
    size_t out = 0;
    for (int i = 0; i < num_macro; i++) {
        rng->a = (uint32_t)i;
        rng->b = 0;
        double vel[3] = {0, 0, 0};
        if (T > 0) { /* sampleV3th, Species.cpp:862-869 */
            double vth = sqrt(2 * ORC_KB * T / mass), v[3];
            for (int c = 0; c < 3; c++) {
                double u1 = orc_rng_draw(rng, 3 * c), u2 = orc_rng_draw(rng, 3 * c + 1);
                double u3 = orc_rng_draw(rng, 3 * c + 2);
                v[c] = vth * (u1 + u2 + u3 - 1.5);
            }
            double sp = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            double theta = 2 * ORC_PI * orc_rng_draw(rng, 9);
            double r = -1.0 + 2 * orc_rng_draw(rng, 10);
            double a = sqrt(1 - r * r);
            vel[0] = sp * r;
            vel[1] = sp * (cos(theta) * a);
            vel[2] = sp * (sin(theta) * a);
        }
        if (plus)
            vel[ax] -= v_drift;
        else
            vel[ax] += v_drift;
        double pos[3];
        pos[ax] = g->x0[ax] + (plus ? L[ax] : 0.0);
        pos[t0] = g->x0[t0] + orc_rng_draw(rng, 11) * L[t0];
        pos[t1] = g->x0[t1] + orc_rng_draw(rng, 12) * L[t1];
        if (!orc_in_bounds(g, pos) || orc_in_object(rects, nrect, pos)) continue;
        if (ef) {
            double lc[3], e[3];
            xtol(g, pos, lc);
            orc_gather3(g, ef, lc, e);
            for (int d = 0; d < 3; d++) vel[d] -= e[d] * (charge / mass) * (0.5 * dt);
        }
        if (out >= max_out) return (size_t)-1;
        double* q = parts + 7 * out;
        for (int d = 0; d < 3; d++) {
            q[d] = pos[d];
            q[3 + d] = vel[d];
        }
        q[6] = mpw0;
        out++;
    }
    return out;
}
