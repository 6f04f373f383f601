/*
 * gample_amd.h -- C-ABI of the MI355X particle-in-cell step.
 *
 * One shared library (libgample_amd.so) built for gfx950.  All device memory
 * (particles, grids, solver hierarchy, network weights) is owned by a
 * context; the host holds opaque handles and receives copies on request.
 * Every call returns an int status (GP_OK = 0; GP_NOT_CONVERGED = 1 is a
 * non-fatal solver status; negative values are errors whose text is given by
 * gp_last_error()).  No C++ exception crosses this boundary.  Calls on one
 * context are ordered on the context's HIP stream and are not re-entrant;
 * separate contexts are independent.
 *
 * Each entry point names the reference interface it replaces (paths are in
 * the reference tree, tnaoki0630/gample).
 */
#ifndef GAMPLE_AMD_H
#define GAMPLE_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GP_OK               0
#define GP_NOT_CONVERGED    1
#define GP_ERR_ARG        (-1)
#define GP_ERR_HIP        (-2)
#define GP_ERR_NOMEM      (-3)
#define GP_ERR_CAPACITY   (-4)
#define GP_ERR_STATE      (-5)
#define GP_ERR_UNSUPPORTED (-6)
#define GP_ERR_NONFINITE  (-7)

/* particle boundary codes (Particle.mm:461-509 function constants 0-3) */
#define GP_PBC_NONE     (-1)
#define GP_PBC_PERIODIC   0
#define GP_PBC_DELETE     1
#define GP_PBC_REFLECT    2
/* field boundary codes (EMField.mm:200-233) */
#define GP_FBC_PERIODIC   0
#define GP_FBC_DIRICHLET  1
#define GP_FBC_NEUMANN    2

typedef struct gp_ctx gp_ctx;

/* Grid, weighting and boundary set-up.  Replaces the parsed
 * ParamForField / BoundaryCondition* structs (Init.h:9-99) as consumed by
 * -[EMField initWithDevice:withParam:withLogger:] (EMField.mm:105-656) and
 * -[Particle initWithDevice:...] (Particle.mm:407-611). */
typedef struct {
    int    ngx, ngy;          /* interior cells (NumberOfGridX/Y) */
    double dx, dy;            /* cell size, cm (GridSizeOfX/Y) */
    int    weight_order;      /* 1, 3 or 5 (WeightingOrder) */
    int    boris_order;       /* 2, 4 or 6 (BorisOrder) */
    int    pbc[4];            /* particle BC Xmin, Xmax, Ymin, Ymax */
    int    fbc[4];            /* field BC Xmin, Xmax, Ymin, Ymax */
    double fbc_value[4];      /* Dirichlet potential, statV */
    double e0[3];             /* uniform initial E, statV/cm (AmplitudeOfE) */
    double b0[3];             /* uniform initial B, G (AmplitudeOfB, static) */
    int    cur_dens;          /* 1: deposit Jx, Jy as well (FlagForCurDens / EMField=2) */
    int    em_model;          /* 1 electrostatic, 2 Darwin post-solve */
    int    device;            /* HIP device ordinal */
} gp_config;

/* Species constants.  Replaces ParamForParticle (Init.h:40-54) and the
 * Particle object's m, q, w members (Particle.mm:420-424). */
typedef struct {
    double q, m, w;           /* statC, g, macro-particle weight */
    int64_t capacity;         /* MaximumParticleNumber */
    double vnorm;             /* velocity scale of the current fixed-point
                                 format, cm/s (rounded up to a power of 2);
                                 0 = choose from vth_hint */
    double vth_hint;          /* thermal speed sqrt(kT/m), cm/s */
} gp_species_desc;

/* Seeded particle generation.  Replaces generationParams and the
 * artificialIonization kernel's initial load (Collision.h:25-41,
 * Collision.mm:63-174, 280-336).  gen_type: 0 uniform, 1 Xsinusoidal,
 * 2 Ysinusoidal, 4 quietStart, 5 Xgaussian, 6 Ygaussian, 7 cosine
 * perturbation 1 + pert_delta*cos(2 pi pert_mode (x-xmin)/(xmax-xmin)). */
typedef struct {
    int   gen_type;
    float xmin, xmax, ymin, ymax;     /* cm */
    float ele_u[3], ele_vth;          /* drift, sqrt(2kT/m) as the reference */
    float ion_u[3], ion_vth;
    float pert_delta; int pert_mode;
} gp_gen;

/* Solver statistics.  Replaces the (iters, error) pair logged by
 * solvePoisson (EMField.mm:866, 1119-1127). */
typedef struct {
    int    iterations;        /* inner preconditioned-CG iterations */
    int    refinements;       /* FP64 true-residual tests of the MG-PCG, plus the
                                 network's FP64 corrections when it ran */
    double relres;            /* ||b - A phi||_2 / ||b||_2 in FP64 */
    double seconds;           /* wall time of the solve up to its last residual check */
    int    used_nn;           /* 1 if the network warm start / FPCG ran */
    int    fell_back;         /* 1 if NN-FPCG fell back to MG-CG */
} gp_solve_stats;

/* per-step diagnostics: the logged scalars of reduce (Particle.mm:990-1001)
 * and solvePoisson (EMField.mm:1119-1127) */
typedef struct {
    double kinetic[8];        /* per species, erg, FP64 */
    double we_ref;            /* sum(E^2) dx dy / 2 over whole arrays (reference form) */
    double we_cgs;            /* E^2/(8 pi) over unique points, erg/cm */
    double joule;             /* sum(E.J) dt dx dy */
    int64_t count[8];         /* live particles per species */
    int64_t absorbed[8][4];   /* particles absorbed per species per wall this step */
    double mean_cathode;      /* statV: mean potential on the hollow-cathode column
                                 (EMField.mm:1063-1071, logged as meanCathode :1122);
                                 0 without a cathode */
} gp_diag;

/* field identifiers for gp_field_get / gp_field_set */
enum {
    GP_F_RHO = 0,    /* double (nx+1)(ny+1), folded total charge density */
    GP_F_EX,         /* float (nx+2)(ny+1) */
    GP_F_EY,         /* float (nx+1)(ny+2) */
    GP_F_EZ,         /* float (nx+1)(ny+1) */
    GP_F_BX,         /* float (nx+1)(ny+2) */
    GP_F_BY,         /* float (nx+2)(ny+1) */
    GP_F_BZ,         /* float (nx+2)(ny+2) */
    GP_F_PHI,        /* double (nx+3)(ny+3) */
    GP_F_JX,         /* double (nx+2)(ny+1) total current (not folded) */
    GP_F_JY,         /* double (nx+1)(ny+2) */
    GP_F_DEL_EX,     /* float, Darwin transverse correction */
    GP_F_DEL_EY,
    GP_F_DEL_BZ,
    GP_F_SOL,        /* double Poisson unknowns (nkx+1)(nky+1) */
    GP_F_CHI,        /* double Darwin potentials on the node grid (nkxv+1)(nkyv+1), */
    GP_F_AX,         /*   nkxv = ngx - [periodic x] (EMField.mm:486-495) */
    GP_F_AY,
    GP_F_RHS,        /* double Poisson right-hand side b of the SPD system (same grid as GP_F_SOL) */
    GP_F_COUNT
};
/* per-species accumulator identifiers for gp_species_grid_get */
enum { GP_G_RHO_ACC = 0, GP_G_JX_ACC, GP_G_JY_ACC };

/* ---- context ---------------------------------------------------------- */
/* Replaces EMField init (EMField.mm:105-656): allocates fields, builds the
 * Poisson/Darwin solver hierarchy. */
int  gp_ctx_create(const gp_config *cfg, gp_ctx **out);
int  gp_ctx_destroy(gp_ctx *ctx);
const char *gp_last_error(void);
int  gp_sync(gp_ctx *ctx);
/* grid geometry: nx = ngx+2ngb, ny, nkx, nky, ikmin, jkmin */
int  gp_dims(const gp_ctx *ctx, int out[6]);

/* ---- species ---------------------------------------------------------- */
/* Replaces -[Particle initWithDevice:withParam:specimen:withLogger:]
 * (Particle.mm:407-611): device SoA store of capacity particles. */
int  gp_species_create(gp_ctx *ctx, const gp_species_desc *d, int *sid);
/* host SoA -> device (replaces the shared-memory particleBuffer writes,
 * Particle.mm:548-597, and FileIO loadProgress, FileIO.mm:295-352) */
int  gp_species_upload(gp_ctx *ctx, int sid, int64_t n, const float *x, const float *y,
                       const float *vx, const float *vy, const float *vz);
/* device -> host copy (replaces reads of particleBuffer contents) */
int  gp_species_download(gp_ctx *ctx, int sid, int64_t cap, int64_t *n, float *x, float *y,
                         float *vx, float *vy, float *vz);
/* particles first .. first+count-1 in the store's current order, clipped to
 * the live count (*n receives the number copied): the reads of the first
 * particles of particleBuffer in outputPhaseSpace (FileIO.mm:4-39) */
int  gp_species_read(gp_ctx *ctx, int sid, int64_t first, int64_t count, int64_t *n, float *x, float *y,
                     float *vx, float *vy, float *vz);
int  gp_species_count(gp_ctx *ctx, int sid, int64_t *n);
/* per-species int64 accumulators (deposit output before scaling) */
int  gp_species_grid_get(gp_ctx *ctx, int sid, int which, int64_t *host, int64_t count);

/* ---- generation ------------------------------------------------------- */
/* Replaces the initial artificialIonization dispatch of
 * -[Collision initWithDevice:...] (Collision.mm:280-336): n electron/ion
 * pairs appended with PCG32 streams seeded by splitmix64(master_seed) in
 * particle-id order starting at first_id.  sid_ion < 0: electrons only. */
int  gp_load_pair(gp_ctx *ctx, int sid_ele, int sid_ion, const gp_gen *g, uint64_t master_seed,
                  int64_t first_id, int64_t n);
/* -[Collision artificialIonization:] / -[Collision hollowCathode:]
 * (Collision.mm:464-566, 568-678; kernel :63-174): addn particles appended
 * to sid_ele (and, unless sid_ion < 0 or gen_type is 3, addn co-located ions
 * to sid_ion) with the region, drift and thermal speeds of g.  Thread id of
 * the call draws from the context's generation stream id: the PCG32 stream
 * of the master seed of the first seeded gp_load_pair (0 without one),
 * continued from the draws the loads and earlier gp_generate calls made on
 * it, as the reference's persistent RNGState array (Collision.mm:300-313,
 * 507-520).  A region of zero width in x or y is a line source.  addn is the
 * caller's (the reference rounds src w^-1 dt stochastically, and the
 * hollow cathode takes the electron current lost at its wall).  deposit != 0
 * also deposits the new particles into the accumulators. */
int  gp_generate(gp_ctx *ctx, int sid_ele, int sid_ion, const gp_gen *g, int64_t addn, int deposit);
/* the hollow-cathode boundary entry of BoundaryConditionForField
 * (EMField.mm:231-232, 1063-1077): after each Poisson solve the mean of the
 * potential unknowns on column (int)(x_cm / dx) is removed along the ramp
 * phi -= mean ((i - ngb - 2) dx) / x_cm; x_cm <= 0 turns it off. */
int  gp_set_cathode(gp_ctx *ctx, double x_cm);
/* Half-Maxwellian flux injection through the x-min wall (no reference
 * counterpart; the reference's nearest analogue is the hollow-cathode line
 * source, Collision.mm:568-678): n particles with vx from the flux-weighted
 * Rayleigh law, vy, vz Gaussian, x = wall + vx*dt*r.  deposit != 0 also
 * deposits them into the current accumulators. */
int  gp_inject_wall(gp_ctx *ctx, int sid, int64_t n, float vth, const float u[3], double dt,
                    uint64_t seed, uint64_t stream, int deposit);

/* ---- step pieces (reference order) ------------------------------------ */
/* -[EMField resetChargeDensity] (EMField.mm:1363-1378) */
int  gp_reset_grids(gp_ctx *ctx);
/* -[Particle integrateChargeDensity:withMoment:withLogger:]
 * (Particle.mm:1004-1040 + kernel :281-389): int64 fixed-point deposit of
 * every live particle of sid into its accumulators (rho, and J if enabled). */
int  gp_deposit(gp_ctx *ctx, int sid);
/* -[EMField solvePoisson:withLogger:] (EMField.mm:745-1136): scale and sum
 * the species accumulators, fold ghosts, solve to tol (relative residual in
 * FP64), reconstruct phi and E = -grad phi, energies.  Clears the
 * accumulators afterwards.  solver: 0 MG-CG, 1 NN warm start + FPCG. */
int  gp_solve_fields(gp_ctx *ctx, double dt, double tol, int maxiter, int solver,
                     gp_solve_stats *st, gp_diag *diag);
/* Solve the Poisson system S x = b for a caller-given right-hand side b
 * (host FP64, (nkx+1)(nky+1) unknowns as GP_F_SOL), x in/out as the initial
 * guess, with the same solver choices as gp_solve_fields.  The context's own
 * potential, fields and warm start are left unchanged.  No reference
 * counterpart (the reference solves only its own rhs, EMField.mm:853-866). */
int  gp_solve_rhs(gp_ctx *ctx, const double *b, double *x, double tol, int maxiter, int solver,
                  gp_solve_stats *st);
/* -[EMField solveVectorPotential:withLogger:] (EMField.mm:1138-1361): the
 * transverse solves for chi, Ax, Ay from the deposited J, then delEx,
 * delEy, delBz; st[0..2] are the three solves' statistics.  Returns
 * GP_NOT_CONVERGED when any of them stopped short of tol. */
int  gp_solve_darwin(gp_ctx *ctx, double dt, double tol, int maxiter, int solver,
                     gp_solve_stats st[3]);
/* statistics of the three solves of the last Darwin solve, by
 * gp_solve_darwin or inside gp_step (no reference counterpart: the
 * reference logs the three AMGCL iteration counts and errors in its
 * solveVectorPotential section, EMField.mm:1349-1360) */
int  gp_darwin_stats(gp_ctx *ctx, gp_solve_stats out[3]);
/* -[Particle update:withEMField:withMom:withLogger:] (Particle.mm:615-672 +
 * kernel :65-278).  deposit != 0 fuses the next step's deposit of the
 * pushed particles into the same kernel (Particle.mm:281-389). */
int  gp_push(gp_ctx *ctx, int sid, double dt, int deposit);
/* -[Particle reduce:] (Particle.mm:893-1002): compacts absorbed particles
 * (holes filled from the tail), per-wall counts, FP64 kinetic energy. */
int  gp_reduce(gp_ctx *ctx, int sid, int64_t counts[4], double *kinetic);
/* gp_reduce without any device allocation (no hipMalloc, hipFree or
 * allocating library call): the radix histograms and scan block sums of the
 * compaction live in a static 12 MiB device array that each GPU holds for
 * the life of the process.  No counterpart in the reference; it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  Same results as
 * gp_reduce.  Fails with GP_ERR_CAPACITY, changing nothing, when the
 * absorbed particles of this call need more scratch than the array holds
 * (at most 2^24 absorbed particles per call are always accepted).  Every
 * call on one GPU shares its array: calls that overlap on one GPU (other
 * contexts' streams, other host threads) race. */
int  gp_reduce_noalloc(gp_ctx *ctx, int sid, int64_t counts[4], double *kinetic);
/* cell/tile sort of the SoA store by a GPU LSD radix sort (no reference
 * counterpart: the reference keeps insertion order).  Particles are grouped
 * by 16x16-cell tile; inside a tile they are interleaved over the 64 LDS
 * bank classes of their cells, each class in its previous order. */
int  gp_sort(gp_ctx *ctx, int sid);
/* drift-triggered sorting in gp_step (no reference counterpart): a species is
 * sorted before its push when the previous step's push gathered more than
 * fraction * n of its particles outside their tile windows; 0 turns it off.
 * Independent of gp_step's fixed sort_interval. */
int  gp_set_sort_drift(gp_ctx *ctx, double fraction);
/* sorts of species sid so far (out[0]) and the drifted-particle count of
 * its last push that gp_step read back (out[1], -1 before any) */
int  gp_sort_stats(gp_ctx *ctx, int sid, int64_t out[2]);
/* -[Moment integrateMoments:withEMField:withLogger:] (Moment.mm:206-270 +
 * kernel :33-135): n, n*v (3), n*vv (6) node moments in FP64,
 * 10*(nx+1)*(ny+1) doubles, unnormalised as the kernel leaves them. */
int  gp_moments(gp_ctx *ctx, int sid, double *host10);

/* ---- whole step ------------------------------------------------------- */
/* One cycle of the main loop (main.mm:154-254) for every species, with the
 * deposit fused into the push: fields from the current accumulators (and the
 * Darwin solves when em_model = 2, statistics by gp_darwin_stats), push +
 * deposit, reduce, and a cell sort every sort_interval cycles (0 = never).
 * st holds the Poisson solve's statistics; GP_NOT_CONVERGED when any solve
 * stopped short of tol. */
int  gp_step(gp_ctx *ctx, double dt, double tol, int maxiter, int solver, int sort_interval,
             int cycle, gp_solve_stats *st, gp_diag *diag);

/* ---- fields ----------------------------------------------------------- */
int  gp_field_get(gp_ctx *ctx, int which, void *host, int64_t bytes);
int  gp_field_set(gp_ctx *ctx, int which, const void *host, int64_t bytes);
/* diagnostics of the current E and J (EMField.mm:1096-1116) */
int  gp_field_energy(gp_ctx *ctx, double dt, double *we_ref, double *we_cgs, double *joule);

/* ---- timing ----------------------------------------------------------- */
/* HIP-event timing on the context stream: enable = 1 starts recording, and
 * gp_kernel_time returns the total ms and launch count since enabling of
 * which = 0 every push (+deposit), 1 every sort, 2 field solves, 3 every
 * reduce, and for species s: 4 + 3 s its push, 5 + 3 s its reduce, 6 + 3 s
 * its sort (the phases main.mm:176-228 times per cycle as update_<s>,
 * reduce_<s> and solvePoisson); 28 the push+deposit kernel alone (every
 * species, without the window reduction and energy sum that follow it). */
int  gp_timing(gp_ctx *ctx, int enable);
int  gp_kernel_time(gp_ctx *ctx, int which, double *total_ms, int64_t *launches);
/* HBM streaming rate of the hot push's memory pattern on this GPU (a ceiling
 * for the push's roofline): reps launches of a kernel that reads the
 * species' x, y, vx, vy and writes four arrays of the same length (32 B per
 * particle) into its scratch store, leaving the particles unchanged; the mean
 * ms per launch.  No counterpart in the reference (a measurement aid). */
int  gp_stream_probe(gp_ctx *ctx, int sid, int reps, double *ms_per_launch);

/* ---- learned preconditioner ------------------------------------------ */
/* Build and train the convolutional warm start / preconditioner on (rho,
 * phi) pairs held by the caller (host arrays, nsamp x (nkx+1)(nky+1)
 * doubles), BF16 MFMA kernels with FP32 accumulate, Adam.  No reference
 * counterpart (the reference solves with AMGCL only, EMField.mm:866). */
int  gp_nn_init(gp_ctx *ctx, int base_channels, uint64_t seed);
int  gp_nn_train(gp_ctx *ctx, const double *rho, const double *phi, int nsamp, int epochs,
                 int batch, float lr, float *loss_out);
int  gp_nn_infer(gp_ctx *ctx, const double *rhs, double *out);
int  gp_nn_save(gp_ctx *ctx, const char *path);
int  gp_nn_load(gp_ctx *ctx, const char *path);
/* Training loss (0 mean square, 1 relative energy norm: the default), the
 * red-black sweeps that follow the network inside the flexible-CG
 * preconditioner (default 2), and the flexible-CG iterations allowed before
 * it falls back to MG-PCG (0, the default: as many as the last plain MG-PCG
 * solve on the same hierarchy needed). */
int  gp_nn_config(gp_ctx *ctx, int loss_kind, int smooth_sweeps, int fcg_cap);
/* Append the Poisson solver's current (b, x) pair (the last solve's
 * right-hand side and solution, device resident) to the training set. */
int  gp_nn_add_pair(gp_ctx *ctx);
/* Append the time-loop increment of the Poisson solver's last two solves,
 * (b_n - b_{n-1}, x_n - x_{n-1}), formed on the device: the residual a solve
 * warm-started from x_{n-1} begins with and the correction it needs.  The
 * first call after gp_nn_init records b_n and x_n only (*added = 0). */
int  gp_nn_add_increment(gp_ctx *ctx, int *added);
/* The hierarchy the training pairs belong to: 0 the Poisson solve (the
 * default), 1 the Darwin solves (em_model = 2).  Changing it drops the pairs
 * held for the other one and keeps the weights, so a network trained on
 * Poisson pairs can be fine-tuned on Darwin ones.  The network serves either
 * hierarchy through gp_solve_fields / gp_solve_darwin / gp_step (solver 1..3). */
int  gp_nn_set_system(gp_ctx *ctx, int sys);
/* Append the time-loop increments of the last Darwin solve's three systems
 * (chi, Ax, Ay: b_n - b_{n-1} and x_n - x_{n-1} of each, on the device) to the
 * Darwin training set; a system's first call only records it.  *added: the
 * pairs appended (0 or 3).  Reference: the systems of EMField.mm:1147-1275. */
int  gp_nn_add_darwin_increments(gp_ctx *ctx, int *added);
/* Flat FP32 parameter vector (get and/or set; count returns its length).
 * Order per layer (in, W1..W6, out): weights [co][3x3 tap][ci], then bias. */
int  gp_nn_params(gp_ctx *ctx, float *get, const float *set, int64_t *count);
/* Network image geometry for the Poisson grid: H, W (padded unknown grid). */
int  gp_nn_dims(gp_ctx *ctx, int *H, int *W);
/* One forward pass on B host images (B x H x W FP32) -> outputs. */
int  gp_nn_forward(gp_ctx *ctx, const float *in, int B, float *out);
/* One forward + backward against targets, no update: the loss and the flat
 * FP32 gradient (gp_nn_params order).  kind 0: mean square over the image;
 * kind 1 (the training loss): energy norm (y - t)' S1 (y - t) per image,
 * mean over the batch, S1 the grid-unit Poisson operator on the unknown
 * grid (zero outside it). */
int  gp_nn_grad(gp_ctx *ctx, const float *in, const float *tgt, int B, int kind, float *loss, float *grad);

#ifdef __cplusplus
}
#endif
#endif
