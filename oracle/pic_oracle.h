/*
 * pic_oracle.h -- FP64 CPU oracle for the PIC/DSMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithms (Misquic/Engineering-degree-in-Plasma-Simulations, ch4/v3 and
 * ch2/v1) used as the checker for the HIP kernels.  Only tests/, the smoke
 * entry point and the cpu_baseline leg of bench.py may load it.  The product
 * library never links or calls anything in this directory.
 *
 * Layout conventions (shared with the device library):
 *   - node fields are flat, i-fastest: u = (k*nj + j)*ni + i
 *     (the reference's solver index, PotentialSolver.cpp:421, Field.h:263-265)
 *   - vector fields are interleaved xyz per node: f[3*u + c]
 *   - cell index c = (k*(nj-1) + j)*(ni-1) + i   (World.cpp:134-137)
 *   - particles are 7 doubles {x,y,z,vx,vy,vz,w} (Species.h:12-29)
 *
 * Every function names the reference lines it restates.  Functions with no
 * check beyond retyping are marked "parity unpinned" in their comment.
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* physical constants, all.h:13-23 (pi truncated as in the reference) */
#define ORC_EPS0 8.85418782e-12
#define ORC_QE 1.602176565e-19
#define ORC_AMU 1.660538921e-27
#define ORC_ME 9.10938215e-31
#define ORC_KB 1.380648e-23
#define ORC_PI 3.141592653

typedef struct {
    int ni, nj, nk;
    double x0[3], xm[3], dx[3], inv_dx[3];
} orc_grid;

typedef struct {
    double c[3];    /* centre */
    double half[3]; /* half side lengths */
    double phi;     /* electrode potential */
} orc_rect;

/* ---- geometry: World.cpp:63-78, 123-137, 201-205, 293-301, 353-367 ---- */
void orc_grid_init(orc_grid* g, int ni, int nj, int nk, const double* x0, const double* xm);
void orc_node_volumes(const orc_grid* g, double* vol);
int orc_in_bounds(const orc_grid* g, const double* pos);
int orc_in_object(const orc_rect* r, int nrect, const double* pos);
void orc_object_id(const orc_grid* g, const orc_rect* r, int nrect, int32_t* object_id,
                   int32_t* node_type, double* phi);
void orc_dirichlet_walls(const orc_grid* g, int32_t* node_type);

/* ---- CIC: Field.h:157-199 (scatter), Field.h:201-232 (gather) ---- */
double orc_gather1(const orc_grid* g, const double* f, const double* lc);
void orc_gather3(const orc_grid* g, const double* f3, const double* lc, double* out);
void orc_scatter1(const orc_grid* g, double* f, const double* lc, double val);

/* ---- particles ---- */
/* Species.cpp:356-399 (serial electron push); stable!=0 keeps survivor order
 * instead of the reference's swap-with-last removal.  Returns survivors. */
size_t orc_push(const orc_grid* g, const double* ef, double* parts, size_t n, double charge,
                double mass, double dt, const orc_rect* rects, int nrect, int stable);
/* Species.cpp:401-416 */
void orc_number_density(const orc_grid* g, const double* parts, size_t n, const double* vol,
                        double* den);
/* World.cpp:193-200: rho += q_s * den_s */
void orc_charge_accumulate(size_t nn, double* rho, double charge, const double* den);
/* Species.cpp:767-776: n, nv(3), nuu, nvv, nww accumulated (not cleared) */
void orc_sample_moments(const orc_grid* g, const double* parts, size_t n, double* n_sum,
                        double* nv_sum, double* nuu, double* nvv, double* nww);
/* Species.cpp:777-804 (T uses m/(2k), reproduced as is) */
void orc_gas_properties(const orc_grid* g, double mass, const double* n_sum, const double* nv_sum,
                        const double* nuu, const double* nvv, const double* nww, double* vel,
                        double* T);
/* Species.cpp:813-819 (macro particle count per cell) */
void orc_macro_count(const orc_grid* g, const double* parts, size_t n, double* count);
/* Species.cpp:905-929: stable per-cell lists as CSR (offsets[ncells+1], index) */
void orc_sort_indexes(const orc_grid* g, const double* parts, size_t n, int32_t* offsets,
                      int32_t* index);
/* Species.cpp:732-763: out = {sum w, px, py, pz, KE, PE} */
void orc_species_sums(const orc_grid* g, const double* parts, size_t n, double mass, double charge,
                      const double* phi, double* out);
/* World.cpp:108-118 */
double orc_field_energy(const orc_grid* g, const double* ef, const double* vol);

/* ---- field solve ---- */
/* PotentialSolver.cpp:354-408 (and ch2/v1/PotentialSolver.cpp:88-150) */
void orc_efield(const orc_grid* g, const double* phi, double* ef);
/* ch2/v1/PotentialSolver.cpp:8-87: interior-only lexicographic SOR,
 * convergence L2 = sqrt(sum R^2 / (ni nj nk)) checked every 25 sweeps. */
int orc_sor_dirichlet(const orc_grid* g, double* phi, const double* rho, int max_it, double tol,
                      double omega, int* iters, double* l2);
/* PotentialSolver.cpp:69-166: v3 GS-SOR with Neumann face copies, object
 * nodes fixed, Boltzmann electrons n0*exp((phi-phi0)/Te0). */
int orc_gs_v3(const orc_grid* g, const int32_t* object_id, double* phi, const double* rho,
              double phi0, double n0, double Te0, int max_it, double tol, double omega, int* iters,
              double* l2);
/* PotentialSolver.cpp:178-241 (Newton), :243-313 (PCG), :315-347 (GS
 * fallback), :414-471 (matrix), Matrix.cpp:99-284.  node_type is the
 * World::node_type flag (2 = DIRICHLET); object_id > 0 is Dirichlet too.
 * axis_swap!=0 reproduces the reference coefficient placement
 * (PotentialSolver.cpp:457-463).  Returns 1 converged, 0 not, -1 when the
 * reference would throw (beta == 0, PotentialSolver.cpp:273-276). */
int orc_nrpcg(const orc_grid* g, const int32_t* object_id, const int32_t* node_type_in,
              double* phi, const double* rho, double phi0, double n0, double Te0, int pcg_max_it,
              double tol, int axis_swap, int* newton_iters, int* pcg_iters_total, double* last_l2);
/* matrix-free 7-point Laplacian residual on Dirichlet-walled interior:
 * r = b - A phi with b = -rho/eps0; returns ||r||_2 and ||b||_2. */
void orc_residual_dirichlet(const orc_grid* g, const double* phi, const double* rho, double* r2,
                            double* b2);
/* dense Gaussian elimination of the Dirichlet-walled Poisson problem; only
 * for tiny grids (interior unknowns <= 1000). */
int orc_dense_poisson(const orc_grid* g, double* phi, const double* rho);

/* ---- random numbers ---- */
typedef struct {
    uint32_t mt[624];
    int mti;
} orc_mt19937;
void orc_mt_seed(orc_mt19937* s, uint32_t seed);
uint32_t orc_mt_next(orc_mt19937* s);
/* libstdc++ uniform_real_distribution<double>(0,1) over mt19937 (Rnd.cpp:11-16) */
double orc_mt_uniform(orc_mt19937* s);
/* Philox4x32-10 (Salmon et al., SC'11): out = bijection(ctr; key) */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_u53(uint32_t a, uint32_t b);

/* draw source used by the loader and by MC_MEX_Ionization: either the
 * reference's sequential mt19937 stream, or the device's counter-based
 * Philox stream addressed by (a, b, slot). */
typedef struct {
    int kind; /* 0 = mt19937, 1 = philox */
    orc_mt19937 mt;
    uint32_t key[2];
    uint32_t tag;
    uint32_t a, b; /* counter words set by the caller (cell/group, particle) */
} orc_rng;
void orc_rng_mt(orc_rng* r, uint32_t seed);
void orc_rng_philox(orc_rng* r, uint32_t seed, uint32_t seq, uint32_t tag);
double orc_rng_draw(orc_rng* r, int slot);

/* ---- loaders: Species.cpp:564-599, :420-434 (addParticle), :855-869 ----
 * mode 0: reference speed sampler (sum of uniforms, Q6 semantics);
 * mode 1: Maxwellian via Box-Muller.  The field ef (may be NULL) is used for
 * the half-step velocity rewind with rewind_dt.  Returns particles written. */
size_t orc_load_box_thermal(const orc_grid* g, const orc_rect* rects, int nrect, const double* ef,
                            double charge, double mass, double rewind_dt, const double* xc,
                            const double* sides, double num_den, double T, double mpw0,
                            orc_rng* rng, int mode, size_t max_out, double* parts);

/* ---- MC_MEX_Ionization: Interactions.cpp:476-980 ---- */
typedef struct {
    double m_neutral, m_electron;
    double E_ion_J;             /* neutral ionization energy [J] */
    double c0, c1, c2;          /* sigma_iz fit, Interactions.cpp:506-508 */
    double B_inc;               /* eV, Interactions.cpp:511 */
    const double* sig_E;        /* elastic table energies [eV] (ascending) */
    const double* sig_v;        /* elastic table cross sections [m^2] */
    int nsig;
    double dt, dV;
    int pool_append;            /* 1: split neutrals join the cell pool (reference) */
} orc_mcc_params;
typedef struct {
    long long collisions, ionizations, groups;
    double wsv_max_step;
} orc_mcc_stats;
double orc_sigma_el(const orc_mcc_params* p, double E);
double orc_sigma_iz(const orc_mcc_params* p, double E);
/* one call of apply_vector_indexes (Interactions.cpp:600-762).
 * neutrals/ions/electrons are growable arrays of 7-double particles with
 * capacities; counts are updated in place.  *wsv_max is W_sigma_v_rel_max.
 * Returns 0, or -1 when a capacity is exceeded. */
int orc_mcc_apply(const orc_grid* g, const orc_mcc_params* p, double* wsv_max, orc_rng* rng,
                  double* neutrals, size_t* nn, size_t cap_n, double* ions, size_t* ni,
                  size_t cap_i, double* electrons, size_t* ne, size_t cap_e, orc_mcc_stats* st);
/* collide() restated for fixed draws (Interactions.cpp:885-947); used by
 * closed-form tests.  Returns ionize flag. */
int orc_mcc_collide(const orc_mcc_params* p, orc_rng* rng, double* vel_neu, double* vel_ele,
                    double* vel_new, double sigma_coll);

/* ---- threaded CPU baseline: v3 structure (Species.cpp:258-355 chunked
 * electron push over a thread pool; Matrix.cpp:117-179 chunked SpMV). ---- */
size_t orc_push_threaded(const orc_grid* g, const double* ef, double* parts, size_t n,
                         double charge, double mass, double dt, int nthreads, int remove);

/* ---- object surfaces (v3/Object.cpp:271-317, v3/Species.cpp:170-256,
 * :835-861) ---- */
void orc_rect_line_intersect(const orc_rect* r, const double* x1, const double* x2, double* t_entry,
                             double* pos, double* n);
/* advanceNonElectron -> advanceNoSputteringSerial: neutrals (charge 0)
 * reflect diffusely, ions are removed at an object and emit
 * int(w/mpw0_n + U) neutrals into neu; see pic_oracle.c for draws and
 * flags.  Returns survivors or (size_t)-1 on neutral overflow. */
size_t orc_push_nonelectron(const orc_grid* g, const double* ef, double* parts, size_t n,
                            double charge, double mass, double dt, const orc_rect* rects, int nrect,
                            double* neu, size_t* n_neu, size_t cap_neu, double mpw0_n,
                            orc_rng* rng, int accept_on_surface, int stable, long long* stats);

/* ---- Species::merge (v3/Species.cpp:987-1145); gpu_mode selects the
 * device's defined variant (see pic_oracle.c). ---- */
int orc_merge(const orc_grid* g, const double* ef, double* parts, size_t* n, size_t cap,
              double charge, double mass, double dt, orc_rng* rng, int gpu_mode, long long* stats);

/* ---- DSMC_MEX (v3/Interactions.cpp:143-284); p2 == NULL: one species ---- */
long long orc_dsmc_apply(const orc_grid* g, double* p1, size_t n1, double mass1, double* p2,
                         size_t n2, double mass2, double mpw0, double dt, double* svmax,
                         orc_rng* rng, long long* groups);

/* ---- ColdBeamSource / WarmBeamSource (v3/Source.cpp:3-185); T <= 0 cold ---- */
size_t orc_source_sample(const orc_grid* g, const orc_rect* rects, int nrect, const double* ef,
                         double charge, double mass, double mpw0, int face, double v_drift,
                         double den, double T, double dt, orc_rng* rng, double* parts,
                         size_t max_out);

#ifdef __cplusplus
}
#endif
#endif
