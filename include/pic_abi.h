/*
 * pic_abi.h -- C ABI of the MI355X (gfx950) PIC/DSMC hot-path library.
 *
 * The reference (Misquic/Engineering-degree-in-Plasma-Simulations, ch4/v3)
 * has no FFI: its boundary is the concrete C++ class API called from
 * main.cpp.  Each entry point below replaces one of those methods; the
 * comment names the reference declaration it stands in for.  The C++ host
 * classes in engineering-degree-in-plasma-simulations_amd/host/ keep the
 * reference's class names and signatures and call these functions.
 *
 * Conventions
 *   - every function returns an int status: PIC_OK (0), a negative PIC_ERR_*
 *     code, or PIC_NOT_CONVERGED (1) from the field solvers; no C++
 *     exception crosses this boundary; pic_last_error() explains failures.
 *   - host memory is caller-owned; device memory is library-owned.
 *   - all work of one context is ordered on one HIP stream owned by the
 *     context; downloads and *_stats outputs are synchronous.
 *   - node fields on the host are flat, i-fastest: u = (k*nj + j)*ni + i
 *     (the reference's solver index, PotentialSolver.cpp:421); vector fields
 *     are interleaved xyz per node.  Cell index c = (k*(nj-1)+j)*(ni-1)+i
 *     (World.cpp:134-137).
 *   - particles cross the boundary as 7 doubles {x,y,z,vx,vy,vz,w}, the
 *     layout of the reference's Particle (Species.h:12-29), so
 *     std::vector<Particle>::data() can be passed directly.
 *   - one context per thread; calls on one context must not overlap.
 */
#ifndef PIC_ABI_H
#define PIC_ABI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIC_OK 0
#define PIC_NOT_CONVERGED 1
#define PIC_ERR_ARG (-1)
#define PIC_ERR_HIP (-2)
#define PIC_ERR_CAPACITY (-3)
#define PIC_ERR_STATE (-4)
#define PIC_ERR_NOMEM (-5)
#define PIC_ERR_NUMERIC (-6)

typedef struct pic_ctx pic_ctx;

/* World(ni,nj,nk, x1,y1,z1, x2,y2,z2): World.cpp:16-21, setExtents :63-78 */
typedef struct {
    int ni, nj, nk;
    double x0[3];
    double xm[3];
} pic_grid_desc;

/* ---- context ---- */
/* World::World (World.cpp:16-21) + computeNodeVolumes (:353-367) */
int pic_create(const pic_grid_desc* grid, pic_ctx** out);
void pic_destroy(pic_ctx* ctx);
const char* pic_last_error(const pic_ctx* ctx);
/* World::setTime (World.cpp:321-324) */
int pic_set_dt(pic_ctx* ctx, double dt);
/* blocks until all queued work of the context is done */
int pic_sync(pic_ctx* ctx);

/* ---- geometry ---- */
/* World::addObject<Rectangle>(centre, phi, sides) + computeObjectID
 * (World.h:110-121, World.cpp:276-292, Object.cpp:166-172, :231-238).
 * Rectangles are absorbing for particles and Dirichlet for the solver. */
int pic_add_rectangle(pic_ctx* ctx, const double centre[3], const double sides[3], double phi);
/* World::dirichletBoundries (World.cpp:263-273) */
int pic_dirichlet_walls(pic_ctx* ctx);
/* World::node_type / object_id public fields (World.h:54-57) as i-fastest
 * int32 arrays (node_type: 0 REGULAR, 1 NEUMANN, 2 DIRICHLET). */
int pic_get_node_flags(pic_ctx* ctx, int32_t* object_id, int32_t* node_type);

/* ---- species ---- */
/* Species(name, mass, charge, world, mpw0) (Species.cpp:30-34); capacity
 * bounds the particle count (births included). */
int pic_species_add(pic_ctx* ctx, double mass, double charge, double mpw0, size_t capacity,
                    int* sid);
/* std::vector<Particle>& getPartRef() (Species.cpp:823-825): replace the
 * particle set (positions outside the box or in an object are dropped at
 * the next sort, as addParticle would, Species.cpp:420-426). */
int pic_particles_upload(pic_ctx* ctx, int sid, const double* aos7, size_t n);
/* const std::vector<Particle>& getConstPartRef() (Species.cpp:820-822).
 * Particles are listed in array order, except for a DSMC neutral species
 * whose births since its last sort are kept behind the sorted part: those
 * are listed cell by cell in the order the next pic_mcc_apply picks them
 * (each cell's particles, scanned in the listed order, form its pick pool). */
int pic_particles_download(pic_ctx* ctx, int sid, double* aos7, size_t capacity, size_t* n);
/* size_t getNumParticles() (Species.cpp:43-45) */
int pic_particles_count(pic_ctx* ctx, int sid, size_t* n);
/* loadParticleBoxThermal(x0, sides, num_den, T) (Species.cpp:564-599) with
 * addParticle's half-step rewind (Species.cpp:427-433).  Draws come from a
 * counter-based Philox4x32-10 stream keyed by seed.  mode 0 reproduces the
 * reference speed sampler (Species.cpp:855-869); mode 1 draws a Maxwellian. */
int pic_load_box_thermal(pic_ctx* ctx, int sid, const double centre[3], const double sides[3],
                         double num_den, double T, uint32_t seed, int mode, size_t* n_loaded);

/* advanceElectrons(dt) / advanceNonElectron(..., dt) (Species.cpp:258-399,
 * :47-77, :170-256): gather E (CIC), kick, drift, remove particles leaving
 * the box; a particle ending in an object follows the species' surface mode
 * (pic_species_set_surface; removed by default).  Removed particles
 * disappear at the next pic_sort. */
int pic_push(pic_ctx* ctx, int sid, double dt);
/* sortIndexes() (Species.cpp:905-929): counting sort of the particle arrays
 * by cell; compacts removed particles; cell offsets stay on the device. */
int pic_sort(pic_ctx* ctx, int sid);
/* pic_push followed by pic_sort (advanceElectrons + sortIndexes,
 * Species.cpp:258-399, :905-929) in two passes over the particles: the
 * first pushes without writing back and builds the sort's histogram, the
 * second pushes again from the same inputs (identical arithmetic) and
 * writes each particle to its sorted position (the first pass keeps each
 * run of equal cells' first slot, so the second needs no atomics).  Same
 * result as the two calls, with about 25% less particle traffic. */
int pic_push_sort(pic_ctx* ctx, int sid, double dt);
/* Behaviour of a species at objects during pic_push / pic_push_sort
 * (advanceNonElectron -> advanceNoSputteringSerial, Species.cpp:170-256;
 * advanceElectrons removes, Species.cpp:333-338):
 *   PIC_SURF_ABSORB   removed (the default for every species);
 *   PIC_SURF_REFLECT  moved to the entry point, diffusely reflected at 300 K
 *                     (sampleReflectedVelocity, Species.cpp:835-861) and
 *                     drifted on for the rest of the step; removed after 20
 *                     bounces (neutrals, Species.cpp:216-219);
 *   PIC_SURF_RECYCLE  removed, emitting int(w / mpw0_t + U) particles of
 *                     species sid_target at the entry point with reflected
 *                     velocities (ions -> neutrals, Species.cpp:220-236).
 * An object hit is decided and traced in FP64 (Rectangle::lineIntersect,
 * Object.cpp:271-317).  Draws: Philox4x32-10 keyed by (seed, the species'
 * push count), counter (particle index, bounce).  Emitted particles are
 * appended to the target; the push fails with PIC_ERR_CAPACITY when they
 * do not fit. */
#define PIC_SURF_ABSORB 0
#define PIC_SURF_REFLECT 1
#define PIC_SURF_RECYCLE 2
int pic_species_set_surface(pic_ctx* ctx, int sid, int mode, int sid_target, uint32_t seed);
/* Species::merge() (Species.cpp:987-1145, called at main.cpp:179-195): in
 * every cell of at least 10 particles the velocities are binned 15 x 15 x 15
 * over the cell's range; each bin of more than two particles becomes two
 * particles of half its weight each, v = mean +- sqrt(variance) per axis
 * (weight, momentum and per-axis energy conserved), placed at two distinct
 * random members' positions; charged species get addParticle's half-step
 * rewind.  The species is sorted and compacted afterwards.  Defined where
 * the reference is not: the range is the cell's own (the source also
 * counts particles[0], :1020), a negative variance is clamped to 0 (the
 * source stops clamping at one below -1e-4, :1091-1098).  Draws: Philox
 * keyed by (seed, the species' merge count), counter (cell, bin).  A cell
 * of more than 8192 particles fails with PIC_ERR_CAPACITY. */
typedef struct {
    long long bins;     /* bins merged */
    long long merged;   /* particles in those bins */
    long long removed;  /* net decrease of the particle count */
} pic_merge_stats;
int pic_merge(pic_ctx* ctx, int sid, uint32_t seed, pic_merge_stats* stats);
/* computeNumberDensity() (Species.cpp:401-416); with PIC_DEPOSIT_MOMENTS
 * also sampleMoments() (Species.cpp:767-776), accumulated until
 * pic_clear_samples (Species.cpp:805-812).  Unsorted particles are sorted
 * first. */
#define PIC_DEPOSIT_DENSITY 1
#define PIC_DEPOSIT_MOMENTS 2
int pic_deposit(pic_ctx* ctx, int sid, int flags);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The density of pic_deposit(ctx, sid, PIC_DEPOSIT_DENSITY),
 * the same kernels and the same result, with the per-cell corner sums in a
 * fixed 64 MiB device array that each GPU holds until the process ends
 * instead of a buffer the library allocates.  Allocates nothing, sorts
 * nothing and does not wait: the species must be sorted (PIC_ERR_STATE
 * otherwise), the grid at most 2^21 cells (PIC_ERR_CAPACITY otherwise),
 * and the context's device current.  Every call on one GPU shares the array
 * (each call writes all it reads), so calls that overlap on one GPU, from
 * contexts on different streams or threads, race. */
int pic_deposit_density_noalloc(pic_ctx* ctx, int sid);
int pic_clear_samples(pic_ctx* ctx, int sid);
/* World::computeChargeDensity (World.cpp:193-200) over all species */
int pic_charge_density(pic_ctx* ctx);

/* ---- field solve ---- */
/* PotentialSolver::solve() (PotentialSolver.cpp:55-67): GS-SOR (:69-166),
 * Newton + PCG (:178-313), plus the learned-preconditioned variant.  With
 * n0 != 0 (Boltzmann electrons) the PCG, mixed and learned methods run the
 * FP64 Newton loop (:178-241) around their inner solve of J y = F,
 * J = A - diag(P); PIC_TOL_REL then stops on the FP64 nonlinear residual,
 * PIC_TOL_REF_RMS on the reference's RMS(y) < 1e-3. */
#define PIC_SOLVER_SOR_RB 0  /* red-black SOR, Boltzmann term, Neumann faces */
#define PIC_SOLVER_PCG 1     /* Newton outer loop + Jacobi PCG, FP64 */
#define PIC_SOLVER_PCG_MIXED 2 /* FP64 residual / Newton + FP32 inner PCG (refinement) */
#define PIC_SOLVER_LEARNED 3 /* U-Net warm start + preconditioner, flexible CG (Dirichlet walls) */
#define PIC_TOL_REF_RMS 0    /* reference criterion: sqrt(sum R^2 / N) < tol */
#define PIC_TOL_REL 1        /* ||b - A phi||_2 / ||b||_2 < tol (regular nodes) */
#define PIC_SOLVE_AXIS_SWAP 1 /* PCG coefficients placed as in PotentialSolver.cpp:457-463 */
/* Start from the linear extrapolation 2 phi_n - phi_(n-1) of the context's
 * last two solutions instead of the current phi (a time loop's warm start;
 * no counterpart in the reference, which starts from the current phi).  The
 * first flagged solve, and the first after phi is uploaded, starts from phi.
 * The solve runs to the same tolerance either way. */
#define PIC_SOLVE_EXTRAPOLATE 2
/* Red-black SOR on the natural node layout, in place.  Without it, a grid
 * without Neumann faces sweeps on a copy of phi, rho and the node types in
 * red-black split layout (each colour of a row contiguous, so a colour's
 * sweep writes whole lines); both give the same values bit for bit (the
 * flag is for comparisons). */
#define PIC_SOLVE_SOR_PLAIN 4
typedef struct {
    int method;
    int tol_kind;
    double tol;
    int max_it;        /* SOR sweeps / PCG iterations per linear solve */
    double omega;      /* SOR weight (reference 1.4, PotentialSolver.h:38) */
    int check_every;   /* SOR/PCG convergence check interval (reference 25 for SOR) */
    double phi0, n0, Te0; /* setReferenceValues (PotentialSolver.cpp:409-413) */
    int flags;
} pic_solver_params;
typedef struct {
    int converged;
    int iterations;      /* SOR sweeps or total inner PCG iterations */
    int newton_iterations;
    double rel_residual; /* ||b - A phi|| / ||b|| over regular nodes, FP64 */
    double ref_l2;       /* the reference's RMS criterion value */
    double seconds;      /* device time of the solve */
} pic_solver_stats;
int pic_poisson(pic_ctx* ctx, const pic_solver_params* params, pic_solver_stats* stats);
/* PotentialSolver::computeEF (PotentialSolver.cpp:354-408) */
int pic_efield(pic_ctx* ctx);

/* ---- field transfer ---- */
#define PIC_FIELD_PHI 0
#define PIC_FIELD_RHO 1
#define PIC_FIELD_EF 2       /* 3 doubles per node */
#define PIC_FIELD_NODE_VOL 3 /* download only */
int pic_field_upload(pic_ctx* ctx, int which, const double* src);
int pic_field_download(pic_ctx* ctx, int which, double* dst);
#define PIC_SFIELD_DEN 0     /* Species::den */
#define PIC_SFIELD_NSUM 1    /* n_sum */
#define PIC_SFIELD_NVSUM 2   /* nv_sum, 3 per node */
#define PIC_SFIELD_NUU 3
#define PIC_SFIELD_NVV 4
#define PIC_SFIELD_NWW 5
#define PIC_SFIELD_MACRO_COUNT 6 /* per cell (computeMacroParticlesCount, Species.cpp:813-819) */
int pic_species_field_download(pic_ctx* ctx, int sid, int which, double* dst);

/* ---- diagnostics ---- */
/* {getMicroCount, getMomentum x,y,z, getKE, getPE} (Species.cpp:732-763),
 * FP64 device reductions */
int pic_diag_sums(pic_ctx* ctx, int sid, double out[6]);
/* World::getPE (World.cpp:108-118) */
int pic_field_energy(pic_ctx* ctx, double* out);

/* ---- DSMC: MC_MEX_Ionization (Interactions.h:99-143) ---- */
typedef struct {
    double E_ion_J;        /* neutral ionization energy [J] */
    double c0, c1, c2;     /* sigma_iz fit (Interactions.cpp:506-508) */
    double B_inc;          /* eV (Interactions.cpp:511) */
    double wsv_max_init;   /* initial W*sigma*v_rel max (Interactions.cpp:534) */
} pic_mcc_desc;
typedef struct {
    long long collisions, ionizations, groups;
    double wsv_max; /* W_sigma_v_rel_max after the call */
    /* split neutrals that did not join their cell's pick pool: beyond 16 in
     * one cell and call they are listed in a 2^20-word arena per MC_MEX
     * object, and only those finding it full stay out (the reference adds
     * every one, :698-700); 0 in every case tested */
    long long pool_overflow;
} pic_mcc_stats;
/* MC_MEX_Ionization(neutrals, ions, electrons, world, table_path) ctor
 * (Interactions.cpp:476-539); the elastic table is passed as arrays. */
int pic_mcc_create(pic_ctx* ctx, int sid_neutral, int sid_ion, int sid_electron,
                   const pic_mcc_desc* desc, const double* sig_E, const double* sig_v, int nsig,
                   int* mcc_id);
/* Interaction::apply(dt) (Interactions.cpp:567-598 -> :600-762).  The
 * neutrals' and electrons' births of earlier calls stay behind their sorted
 * parts, indexed by cell (joined with them once they pass 0.75 of them;
 * after any other change a species is sorted whole).  A cell's pick pool
 * holds its neutrals and, as in the reference, the neutrals split off
 * earlier in the same call (:698-700).  Draws: Philox4x32-10 keyed by
 * (seed, step), counter (cell, group). */
int pic_mcc_apply(pic_ctx* ctx, int mcc_id, double dt, uint32_t seed, uint32_t step,
                  pic_mcc_stats* stats);

/* ---- DSMC_MEX (Interactions.h:43-66): equal-weight DSMC with Bird's VHS
 * cross section within one species or between two of equal mpw0 ---- */
/* DSMC_MEX(species, world) / DSMC_MEX(species1, species2, world)
 * (Interactions.cpp:143-172): sid2 == sid1 for one species; two species
 * must have the same mpw0 (the constructor's guard).  sigma v_rel max
 * starts at 1e-14 (Interactions.h:60). */
int pic_dsmc_create(pic_ctx* ctx, int sid1, int sid2, int* dsmc_id);
typedef struct {
    long long collisions, groups;
    double sigma_v_rel_max; /* after the call */
} pic_dsmc_stats;
/* DSMC_MEX::apply(dt) (Interactions.cpp:177-259): per cell (int)(0.5 N1 N2
 * mpw0 svmax dt / dV + 0.5) random pairs (distinct within one species),
 * accepted with P = sigma v_rel / svmax, isotropic scattering in the centre
 * of mass; svmax becomes the step's maximum when anything collided.  One
 * thread per cell over the cell-sorted species (sorted first when needed).
 * Draws: Philox keyed (seed, step), counter (cell, group). */
int pic_dsmc_apply(pic_ctx* ctx, int dsmc_id, double dt, uint32_t seed, uint32_t step,
                   pic_dsmc_stats* stats);

/* ---- sources (Source.h / Source.cpp:3-185) ---- */
/* ColdBeamSource (T <= 0) / WarmBeamSource(species, world, v_drift, den,
 * [T,] face) (Source.cpp:3-28, :32-37, :105-109): face 0 x-, 1 x+, 2 y-, 3 y+,
 * 4 z-, 5 z+ (inletName2Index, World.cpp:369-390).  The face's nodes become
 * Dirichlet at phi = 0 (World::addInlet, World.cpp:206-262); each sample
 * injects (int)(den v_drift A dt / mpw0 + U) particles, dt the context's at
 * creation. */
int pic_source_create(pic_ctx* ctx, int sid, int face, double v_drift, double den, double T,
                      int* source_id);
/* Source::sample() (Source.cpp:100-103, :181-185): positions uniform on the
 * face (a + face one cell inside, :47-48), velocity v_drift along the inward
 * normal plus a warm beam's sampleV3th(T), addParticle's rejection and
 * half-step rewind (Species.cpp:420-434); the particles are appended.
 * Draws: Philox keyed (seed, the source's call count), counter (particle,
 * 0); the count from counter (~0, ~0). */
int pic_source_sample(pic_ctx* ctx, int source_id, uint32_t seed, size_t* n_added);

/* ---- learned warm start / preconditioner (no counterpart in the reference) ----
 * A 2..5-level 3-D conv U-Net, BF16 weights on MFMA, FP32 accumulation, on
 * the node grid coarsened by `coarsen` (1: one voxel per node; 2: one voxel
 * per 2x2x2 aggregate of nodes).  Node counts must be divisible by
 * coarsen * 2^(levels-1).  Input channels per voxel: {charge density in grid
 * units, fraction of non-regular (Dirichlet) nodes}; output: the potential
 * of the unit-spacing problem laplacian(phi) = -rho, phi = 0 on Dirichlet
 * nodes, averaged over the aggregate's regular nodes.  With coarsen 2 the
 * network is the coarse solve of the learned solver's two-level
 * preconditioner.  Host arrays are channels-last, voxel order x-fastest. */
#define PIC_UNET_MAX_LEVELS 5
typedef struct {
    int levels;                         /* 2..5 */
    int channels[PIC_UNET_MAX_LEVELS];  /* multiples of 16: level 0 <= 32, others <= 64 */
    int coarsen;                        /* 1 or 2 */
    uint32_t seed;                      /* weight initialisation */
} pic_unet_desc;
/* Training pairs: a residual t (Gaussian blobs plus white noise), the part
 * res = t - A S(t) left by the pre-smoother S, restricted to the network grid
 * (input), and the restricted exact error A^-1 res from the plain solver
 * (target).  Loss per sample: 0.5 |y - target|^2 / (N (alpha ms + 1 - alpha))
 * over the N voxels holding a regular node (alpha = 1: relative error). */
typedef struct {
    int samples;          /* training pairs generated by the plain solver */
    int epochs;
    int batch;            /* samples per Adam step */
    double lr;
    uint32_t seed;
    int blobs_min, blobs_max;     /* Gaussian blobs per residual */
    double sigma_min, sigma_max;  /* blob widths in node spacings */
    double noise;                 /* maximum white-noise RMS (uniform per sample) */
    double rel_weight;            /* alpha in [0, 1): sample weight 1 / (alpha ms + 1 - alpha),
                                     ms = the target's mean square over the set mean */
} pic_unet_train_desc;
typedef struct {
    double loss_first, loss_last;  /* mean relative loss of the first and last epoch */
    double seconds_data, seconds_train;
    int steps;
} pic_unet_train_stats;
int pic_unet_create(pic_ctx* ctx, const pic_unet_desc* desc);
int pic_unet_param_count(pic_ctx* ctx, size_t* n);
int pic_unet_get_params(pic_ctx* ctx, float* dst, size_t n);
int pic_unet_set_params(pic_ctx* ctx, const float* src, size_t n);
/* in2: 2 floats per network voxel; out: 1 float per network voxel */
int pic_unet_forward(pic_ctx* ctx, const float* in2, float* out);
/* loss = 0.5 mean over voxels holding a regular node (mask channel < 1) of
 * (y - target)^2, and its gradient */
int pic_unet_loss_grad(pic_ctx* ctx, const float* in2, const float* target, double* loss,
                       float* grad);
int pic_unet_adam_step(pic_ctx* ctx, const float* grad, float lr);
/* generate pairs with the plain solver on the device and train (Adam) */
int pic_unet_train(pic_ctx* ctx, const pic_unet_train_desc* desc, pic_unet_train_stats* stats);
/* z = M(r): the learned solver's preconditioner alone on host vectors of one
 * double per node (three red-black Gauss-Seidel sweeps from z = 0, the
 * network's coarse correction of the remaining residual, three reverse-order
 * sweeps; the sweeps after the first and z are computed in FP64 and stored in
 * FP32). */
int pic_learned_precondition(pic_ctx* ctx, const double* r, double* z);

/* ---- profiling: HIP events around every launch of the context ---- */
#define PIC_K_PUSH 0
#define PIC_K_SORT 1
#define PIC_K_DEPOSIT 2
#define PIC_K_RHO 3
#define PIC_K_EFIELD 4
#define PIC_K_SOR 5
#define PIC_K_PCG 6
#define PIC_K_MCC 7
#define PIC_K_LOAD 8
#define PIC_K_UNET 9
#define PIC_K_SCATTER 10 /* the sort's scatter pass (PIC_K_SORT: histogram + scan) */
#define PIC_K_COUNT 11
int pic_profile_enable(pic_ctx* ctx, int on);
/* total milliseconds and launch count per category since the last reset */
int pic_profile_read(pic_ctx* ctx, double ms[PIC_K_COUNT], long long launches[PIC_K_COUNT]);
int pic_profile_reset(pic_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
